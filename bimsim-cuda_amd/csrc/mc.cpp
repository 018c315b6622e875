// Monte-Carlo plane-wave directions (mcSampleNA, reference montecarlo.cpp:
// 9-71), host side like the reference.  With BIMSIM_RNG_GLIBC the draws are
// the glibc rand() stream after srand(seed), obtained through random_r on a
// private TYPE_3 state (rand() is random() on a 128-byte TYPE_3 state), so the
// caller's global rand() state is left alone; the arithmetic is float like
// the reference's ptype.  BIMSIM_RNG_SPLITMIX64 draws uniform doubles from
// SplitMix64(seed) and evaluates in double.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "../../include/bimsim_hip.h"

namespace {

const float kPIf = 3.14159f;

// rotation taking +z to k: quaternion about z x k (montecarlo.cpp:23-40,
// rts/math/quaternion.h:41-112), identity when k = z
template <class T>
void rotation_to(const T k[3], T M[3][3]) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) M[a][b] = a == b ? (T)1 : (T)0;
    const T ca = (((T)0 + k[0] * (T)0) + k[1] * (T)0) + k[2] * (T)1;
    if (ca == (T)1.0) return;
    T ax[3] = {(T)0 * k[2] - (T)1 * k[1], (T)1 * k[0] - (T)0 * k[2], (T)0 * k[1] - (T)0 * k[0]};
    const T l = std::sqrt((((T)0 + ax[0] * ax[0]) + ax[1] * ax[1]) + ax[2] * ax[2]);
    for (int i = 0; i < 3; ++i) ax[i] = ax[i] / l;
    const T angle = std::acos(ca);
    const T w = (T)std::cos(angle / 2.0);
    const T x = ax[0] * (T)std::sin(angle / 2.0);
    const T y = ax[1] * (T)std::sin(angle / 2.0);
    const T z = ax[2] * (T)std::sin(angle / 2.0);
    const T x2 = x + x, y2 = y + y, z2 = z + z;
    const T xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
    const T wx = w * x2, wy = w * y2, wz = w * z2;
    M[0][0] = (T)1.0 - (yy + zz);
    M[0][1] = xy - wz;
    M[0][2] = xz + wy;
    M[1][0] = xy + wz;
    M[1][1] = (T)1.0 - (xx + zz);
    M[1][2] = yz - wx;
    M[2][0] = xz - wy;
    M[2][1] = yz + wx;
    M[2][2] = (T)1.0 - (xx + yy);
}

template <class T>
void rotate(const T M[3][3], const T c[3], double* out) {
    for (int r = 0; r < 3; ++r) {
        T acc = 0;
        for (int cc = 0; cc < 3; ++cc) acc += M[r][cc] * c[cc];
        out[r] = acc;
    }
}

struct SplitMix64 {
    uint64_t s;
    double uniform() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

extern "C" int bimsim_mc_sample_na(double* out, int N, const double k[3], double NAin,
                                   double NAout, int rng, unsigned long long seed) {
    if (N < 0 || (N > 0 && !out) || !k) return BIMSIM_E_INVALID;
    if (!(NAin >= 0.0 && NAout >= NAin && NAout <= 1.0)) return BIMSIM_E_INVALID;
    if (rng == BIMSIM_RNG_GLIBC) {
        char state[128];
        struct random_data rd;
        std::memset(&rd, 0, sizeof(rd));
        std::memset(state, 0, sizeof(state));
        if (initstate_r((unsigned)seed, state, sizeof(state), &rd) != 0) return BIMSIM_E_INVALID;
        auto draw = [&]() {
            int32_t v = 0;
            random_r(&rd, &v);
            return (float)v / (float)RAND_MAX;
        };
        const float kk[3] = {(float)k[0], (float)k[1], (float)k[2]};
        float M[3][3];
        rotation_to(kk, M);
        const float inZ = std::cos(std::asin((float)NAin));
        const float outZ = std::cos(std::asin((float)NAout));
        const float rangeZ = inZ - outZ;
        for (int i = 0; i < N; ++i) {
            const float z = draw() * rangeZ + outZ;
            const float theta = draw() * 2 * (float)kPIf;
            const float phi = std::acos(z);
            const float c[3] = {1 * std::cos(theta) * std::sin(phi),
                                1 * std::sin(theta) * std::sin(phi), 1 * std::cos(phi)};
            rotate(M, c, out + 3 * i);
        }
        return BIMSIM_OK;
    }
    if (rng == BIMSIM_RNG_SPLITMIX64) {
        double M[3][3];
        rotation_to(k, M);
        const double inZ = std::cos(std::asin(NAin)), outZ = std::cos(std::asin(NAout));
        const double rangeZ = inZ - outZ;
        SplitMix64 g{seed};
        for (int i = 0; i < N; ++i) {
            const double z = g.uniform() * rangeZ + outZ;
            const double theta = g.uniform() * 2 * (double)kPIf;
            const double phi = std::acos(z);
            const double c[3] = {std::cos(theta) * std::sin(phi), std::sin(theta) * std::sin(phi),
                                 std::cos(phi)};
            rotate(M, c, out + 3 * i);
        }
        return BIMSIM_OK;
    }
    return BIMSIM_E_INVALID;
}
