// fft_kernels.hip -- public launchers of the LDS FFT passes: each length goes to the
// translation unit (fft_g*.hip) that instantiates its kernels.
#include "fft_passes.hpp"

namespace hp {

#define HP_GROUPS(M) M(g0) M(g1) M(g2) M(g3) M(g4)
#define HP_DECL(G)                                                                                               \
    bool launch_rows_fwd_##G(int, hipStream_t, const float*, float2*, const int*, int, const Geo&, const float2*,  \
                             const FftVariant&);                                                                 \
    bool launch_cols_##G(int, hipStream_t, int, float2*, int, const Geo&, const float2*, const FftVariant&);      \
    bool launch_spectral_##G(int, int, hipStream_t, const float2*, float2*, const float*, const SpecWork*, int,   \
                             const int*, const Geo&, const float2*, const FftVariant&);                          \
    bool launch_rows_inv_##G(int, hipStream_t, const float2*, float*, const TgtWork*, int, const int*,            \
                             const GrowthParams*, const Geo&, const float2*, const FftVariant&);                 \
    bool launch_rows_inv_conv_##G(int, hipStream_t, const float2*, float*, const TgtWork*, int, const int*,       \
                                  const Geo&, const float2*, const FftVariant&);                                \
    bool launch_batch_rows_fwd_##G(int, hipStream_t, const MemberDesc*, const int4*, int, int, const FftVariant&); \
    bool launch_batch_spectral_##G(int, hipStream_t, const MemberDesc*, const int4*, int, const FftVariant&);    \
    bool launch_batch_rows_inv_##G(int, hipStream_t, const MemberDesc*, const int4*, int, bool, int,             \
                                   const FftVariant&);
HP_GROUPS(HP_DECL)
#undef HP_DECL

bool fft_supported(int L) {
#define X_(n) if (L == n) return true;
    HP_FFT_G0(X_) HP_FFT_G1(X_) HP_FFT_G2(X_) HP_FFT_G3(X_) HP_FFT_G4(X_)
#undef X_
    return false;
}

int spectral_block_lines(int L, const FftVariant& v) {
#define X_(n) if (L == n) return v.cs == 2 ? Cfg<n, 2>::CB : Cfg<n, 1>::CB;
    HP_FFT_G0(X_) HP_FFT_G1(X_) HP_FFT_G2(X_) HP_FFT_G3(X_) HP_FFT_G4(X_)
#undef X_
    return 0;
}

#define HP_ANY(call) (call##_g0 ARGS || call##_g1 ARGS || call##_g2 ARGS || call##_g3 ARGS || call##_g4 ARGS)

bool launch_rows_fwd(int L, hipStream_t st, const float* A, float2* X, const int* src_chan, int nsrc, const Geo& g,
                     const float2* tw, const FftVariant& v) {
#define ARGS (L, st, A, X, src_chan, nsrc, g, tw, v)
    return HP_ANY(launch_rows_fwd);
#undef ARGS
}

bool launch_cols(int L, hipStream_t st, int dir, float2* X, int nspec, const Geo& g, const float2* tw,
                 const FftVariant& v) {
#define ARGS (L, st, dir, X, nspec, g, tw, v)
    return HP_ANY(launch_cols);
#undef ARGS
}

bool launch_spectral(int L, int dim, hipStream_t st, const float2* X, float2* Y, const float* Kq,
                     const SpecWork* work, int nwork, const int* klist, const Geo& g, const float2* tw,
                     const FftVariant& v) {
#define ARGS (L, dim, st, X, Y, Kq, work, nwork, klist, g, tw, v)
    return HP_ANY(launch_spectral);
#undef ARGS
}

bool launch_rows_inv(int L, hipStream_t st, const float2* Y, float* U, const TgtWork* work, int nwork,
                     const int* klist, const GrowthParams* gp, const Geo& g, const float2* tw, const FftVariant& v) {
#define ARGS (L, st, Y, U, work, nwork, klist, gp, g, tw, v)
    return HP_ANY(launch_rows_inv);
#undef ARGS
}

bool launch_rows_inv_conv(int L, hipStream_t st, const float2* Y, float* conv, const TgtWork* work, int nwork,
                          const int* klist, const Geo& g, const float2* tw, const FftVariant& v) {
#define ARGS (L, st, Y, conv, work, nwork, klist, g, tw, v)
    return HP_ANY(launch_rows_inv_conv);
#undef ARGS
}

bool launch_batch_rows_fwd(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur,
                           const FftVariant& v) {
#define ARGS (L, st, mem, items, nitems, cur, v)
    return HP_ANY(launch_batch_rows_fwd);
#undef ARGS
}

bool launch_batch_spectral(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems,
                           const FftVariant& v) {
#define ARGS (L, st, mem, items, nitems, v)
    return HP_ANY(launch_batch_spectral);
#undef ARGS
}

bool launch_batch_rows_inv(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, bool conv,
                           int cslot, const FftVariant& v) {
#define ARGS (L, st, mem, items, nitems, conv, cslot, v)
    return HP_ANY(launch_batch_rows_inv);
#undef ARGS
}

}  // namespace hp
