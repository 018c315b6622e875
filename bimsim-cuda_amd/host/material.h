// rts::material<T> (reference rts/optics/material.h:194-665): a refractive
// index n(lambda) = dispersion(lambda) + n0.  A constant material stores one
// dispersion point (lambda, i k) and n0 = n; a material file stores
// (lambda, n + i k) rows and n0 = 0.  Between rows the dispersion is
// interpolated with the reference's weights, which are reversed
// (rts/math/function.h:69-74: riMin*a + riMax*(1-a)); outside the rows it is
// held at the end values.
#pragma once

#include <complex>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dataTypes.h"

namespace rts {

template <class T>
class material {
   public:
    struct point {
        T lambda;
        std::complex<T> n;
    };

    material(T lambda = 1.0, T n = 1.4, T k = 0.0) : n0(n) { insert(lambda, std::complex<T>(0.0, k)); }

    // [lambda n k] rows, optional "#...[format]" first line with the fields
    // microns / wavenumber / n / k / A (rts/optics/material.h:448-541)
    explicit material(const std::string& filename, const std::string& format = "", T scaleA = 1.0);

    static material fromString(const std::string& text, const std::string& format = "", T scaleA = 1.0);

    std::complex<T> operator()(T l) const { return linear(l) + n0; }
    std::complex<T> getN(T l) const { return (*this)(l); }
    T getN0() const { return n0; }
    const std::string& getName() const { return name; }
    size_t size() const { return f.size(); }
    const point& entry(size_t i) const { return f[i]; }

    // insert keeping lambda ascending (after equal lambdas, like function::insert)
    void insert(T lambda, std::complex<T> y) {
        size_t i = 0;
        while (i < f.size() && !(f[i].lambda > lambda)) ++i;
        f.insert(f.begin() + i, point{lambda, y});
    }

   private:
    std::vector<point> f;
    T n0 = 0;
    std::string name;

    std::complex<T> linear(T x) const {
        if (f.empty()) return std::complex<T>(0, 0);
        size_t i = 0;
        while (i < f.size() && !(f[i].lambda > x)) ++i;
        if (i == f.size()) return f.back().n;
        if (i == 0) return f.front().n;
        const T xMax = f[i].lambda, xMin = f[i - 1].lambda;
        const T a = (x - xMin) / (xMax - xMin);
        return f[i - 1].n * a + f[i].n * ((T)1.0 - a);
    }
};

template <class T>
material<T> material<T>::fromString(const std::string& text, const std::string& format_in, T scaleA) {
    material<T> m;
    m.f.clear();
    m.n0 = 0;
    std::stringstream ss(text);
    std::string line, format = format_in;
    if (format.empty()) {
        format = "microns,n,k";
        if (ss.peek() == '#') {
            std::getline(ss, line);
            const size_t a = line.find('['), b = line.find(']');
            if (a != std::string::npos && b != std::string::npos && b > a)
                format = line.substr(a + 1, b - a - 1);
        }
    }
    // field list of the format string (entryType, material.h:53-90)
    enum Field { kMicrons, kWavenumber, kN, kK, kA, kIgnore };
    std::vector<Field> fields;
    {
        std::string rest = format;
        while (true) {
            const size_t e = rest.find(',');
            const std::string tok = rest.substr(0, e);
            fields.push_back(tok == "microns" ? kMicrons : tok == "wavenumber" ? kWavenumber
                             : tok == "n" ? kN : tok == "k" ? kK : tok == "A" ? kA : kIgnore);
            if (e == std::string::npos) break;
            rest = rest.substr(e + 1);
        }
    }
    while (std::getline(ss, line)) {
        // comment rows are skipped; blank rows too (the reference's parser
        // does not terminate on them)
        if (line.empty() || line[0] == '#' || line.find_first_of("0123456789") == std::string::npos)
            continue;
        std::stringstream ls(line);
        T lambda = 0;
        std::complex<T> n(0, 0);
        for (Field fld : fields) {
            // skip to the next digit: signs and leading dots are dropped as in
            // entryType::inputEntry (material.h:115-123)
            while (ls.peek() != EOF && (ls.peek() < '0' || ls.peek() > '9')) ls.ignore();
            // values are read as float, the reference's ptype
            float fv = 0;
            if (!(ls >> fv)) break;
            const T val = fv;
            switch (fld) {
                case kMicrons: lambda = val; break;
                case kWavenumber: lambda = (T)10000.0 / val; break;
                case kN: n.real(val); break;
                case kK: n.imag(val); break;
                case kA: n.imag((val * scaleA * lambda) / (4 * PI)); break;
                default: break;
            }
        }
        m.insert(lambda, n);
    }
    return m;
}

template <class T>
material<T>::material(const std::string& filename, const std::string& format, T scaleA) {
    std::ifstream ifs(filename.c_str());
    if (!ifs.is_open()) throw std::runtime_error("Material Error -- file not found: " + filename);
    const std::string text((std::istreambuf_iterator<char>(ifs)), std::istreambuf_iterator<char>());
    *this = fromString(text, format, scaleA);
    name = filename;
}

}  // namespace rts
