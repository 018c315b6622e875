// planewave<T> (reference planewave.h:9-61): wave vector k and a field E
// made perpendicular to k.
#pragma once

#include <ostream>
#include <sstream>
#include <string>

#include "dataTypes.h"

template <class T>
class planewave {
    rts::vector<T, 3> k;
    rts::vector<T, 3> E;

   public:
    // an x-polarised wave along +z
    planewave() : k(0, 0, 1), E(1, 0, 0) {}

    // E is replaced by its component perpendicular to k
    planewave(rts::vector<T, 3> k_vec, rts::vector<T, 3> E0) : k(k_vec) {
        const rts::vector<T, 3> s = E0.cross(k);
        rts::vector<T, 3> E_hat;
        if (s.len() == 0) E_hat = rts::vector<T, 3>(0, 0, 0);
        else E_hat = (s.cross(k)).norm();
        E = E_hat * (E_hat.dot(E0));
    }

    // the reference leaves refraction unimplemented (planewave.h:40-46)
    void refract(rts::vector<T, 3> v) { (void)v; }

    const rts::vector<T, 3>& kvec() const { return k; }
    const rts::vector<T, 3>& Evec() const { return E; }

    std::string toStr() {
        std::stringstream ss;
        ss << "k = " << k << std::endl;
        ss << "E = " << E << std::endl;
        return ss.str();
    }
};

template <typename T>
std::ostream& operator<<(std::ostream& os, planewave<T> p) {
    return os << p.toStr();
}
