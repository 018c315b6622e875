// The harmonic pixel kernel of one order bucket (HARM_L, set by the build:
// this file is compiled three times per bucket) in its register-budget
// variants: HARM_PART 0 the variants that hold all order sums at once,
// HARM_PART 1 the chunked planar variants, HARM_PART 2 the chunked planar
// variants with the order sums on the matrix cores.
#include "us_harm.h"

#if !defined(HARM_L) || !defined(HARM_PART)
#error "us_harm.hip is compiled per order bucket with -DHARM_L=<orders> -DHARM_PART=<0|1|2>"
#endif

namespace bimsim {
namespace {

#if HARM_PART == 0
template <int L, int W>
void launch_variant(const HarmLaunch& a) {
    // buckets whose spheres are all planar take the kernel without the
    // general branch (fewer registers live)
    if (a.planar)
        hipLaunchKernelGGL((harm::series_harm_kernel<L, true, W, 0>), a.grid,
                           dim3(harm::kHarmThreads), 0, a.stream, a.inst, a.sph, a.grp, a.items,
                           a.n_items, a.rec, a.rec_in, a.cst, a.perm, a.U, a.counter);
    else
        hipLaunchKernelGGL((harm::series_harm_kernel<L, false, W, 0>), a.grid,
                           dim3(harm::kHarmThreads), 0, a.stream, a.inst, a.sph, a.grp, a.items,
                           a.n_items, a.rec, a.rec_in, a.cst, a.perm, a.U, a.counter);
}

#else
// the chunked planar form (planar buckets only), its sums formed by VALU
// FMAs or on the matrix cores
template <int L, int W, int C, bool kMfma>
int launch_chunked(const HarmLaunch& a) {
    if constexpr (W > 0 && C > 0) {
        if (!a.planar) return -1;
        hipLaunchKernelGGL((harm::series_harm_kernel<L, true, W, C, kMfma>), a.grid,
                           dim3(harm::kHarmThreads), 0, a.stream, a.inst, a.sph, a.grp, a.items,
                           a.n_items, a.rec, a.rec_in, a.cst, a.perm, a.U, a.counter);
        return 0;
    }
    return -1;
}
#endif

}  // namespace

#if HARM_PART == 0
template <int L>
int harm_launch_full(const HarmLaunch& a, int variant) {
    constexpr int w0 = harm_variant_waves(L, 0);
    constexpr int w1 = harm_variant_waves(L, 1);
    constexpr int w2 = harm_variant_waves(L, 2);
    switch (variant) {
        case 0:
            launch_variant<L, w0>(a);
            return 0;
        case 1:
            launch_variant<L, w1>(a);
            return 0;
        case 2:
            if constexpr (w2 > 0) {
                launch_variant<L, w2>(a);
                return 0;
            }
            return -1;
        default:
            return -1;
    }
}

template int harm_launch_full<HARM_L>(const HarmLaunch& a, int variant);
#elif HARM_PART == 1
template <int L>
int harm_launch_chunked(const HarmLaunch& a, int variant) {
    switch (variant) {
        case 3:
            return launch_chunked<L, harm_variant_waves(L, 3), harm_variant_chunk(L, 3), false>(a);
        case 4:
            return launch_chunked<L, harm_variant_waves(L, 4), harm_variant_chunk(L, 4), false>(a);
        case 5:
            return launch_chunked<L, harm_variant_waves(L, 5), harm_variant_chunk(L, 5), false>(a);
        default:
            return -1;
    }
}

template int harm_launch_chunked<HARM_L>(const HarmLaunch& a, int variant);
#else
template <int L>
int harm_launch_mfma(const HarmLaunch& a, int variant) {
    switch (variant) {
        case 6:
            return launch_chunked<L, harm_variant_waves(L, 6), harm_variant_chunk(L, 6), true>(a);
        case 7:
            return launch_chunked<L, harm_variant_waves(L, 7), harm_variant_chunk(L, 7), true>(a);
        case 8:
            return launch_chunked<L, harm_variant_waves(L, 8), harm_variant_chunk(L, 8), true>(a);
        default:
            return -1;
    }
}

template int harm_launch_mfma<HARM_L>(const HarmLaunch& a, int variant);
#endif

}  // namespace bimsim
