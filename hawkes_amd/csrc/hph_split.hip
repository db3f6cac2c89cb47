// Split passes for spatially extended data: the background pass on the culled order (bounding-box
// skips) and the kernel that turns its partials into per-event sums; the excitation sums come from
// excitation_band_kernel (hph_kernels.hip).
#include "hph_sym_device.h"

#include <stdexcept>

namespace hph {
namespace kern {

void launch_sym_split(const SymLaunch &L, const double *params, bool grad, bool fp32, double *rowpart,
                      double *colpart, hipStream_t s) {
    if (L.fast != 0 && L.fast != 1) throw std::invalid_argument("split passes take one form for the whole pass");
    if (L.ktab != 0) throw std::invalid_argument("split passes run the table exponential with its ldexp");
    if (L.fast != 0 && L.b2tab != (fp32 ? 0 : 1)) throw std::invalid_argument("base-2 units do not match the precision");
    const int v = (fp32 ? 2 : 0) + L.fast;
    if (L.pm == kPassCulledB) {
        switch (v) {
        case 0: launch_sym_p<false, 0, kPassCulledB>(L, params, grad, rowpart, colpart, s); break;
        case 1: launch_sym_p<false, 1, kPassCulledB>(L, params, grad, rowpart, colpart, s); break;
        case 2: launch_sym_p<true, 0, kPassCulledB>(L, params, grad, rowpart, colpart, s); break;
        default: launch_sym_p<true, 1, kPassCulledB>(L, params, grad, rowpart, colpart, s); break;
        }
    } else {
        throw std::invalid_argument("unknown pass mode");
    }
}

// Background sums per event of a kPassCulledB pass over dataset 0: for culled position r, its
// row-side chunks, its diagonal block and the column side of every later block, in a fixed
// order, written to its time-sorted position perm[r]. One block of 256 positions per workgroup.
template <bool GRAD>
__global__ __launch_bounds__(kRowsPerBlock) void sym_bsum_kernel(SymLaunch L, const double *__restrict__ params,
                                                                 const double *__restrict__ rowpart,
                                                                 const double *__restrict__ colpart,
                                                                 double *__restrict__ bsum) {
    constexpr int NR = GRAD ? 6 : 2;
    constexpr int NC = GRAD ? 3 : 1;
    constexpr int QS = GRAD ? 3 : 1;
    __shared__ int qlist[kQChunk], rlist[kQChunk];
    __shared__ int qcount, rcount;
    const int4 it = L.fin_item[blockIdx.x];
    if (it.x != 0) return; // uniform: the whole workgroup
    const SymData &ds = L.data[0];
    const int k = it.y;
    const int r = it.z * kRowsPerBlock + threadIdx.x;
    const int R = L.R;
    // column side: the workgroup's 256 positions are column tile it.z of every later block; the
    // first kQChunk later blocks that wrote it are added with the row partials, the rest after
    const int Pw = it.z * kRowsPerBlock / R;
    const double *cbase = colpart + ds.col0 + (size_t)k * ds.col_kstride;
    const unsigned char *ftile = L.colflag + ds.flag0 + (size_t)k * ds.flag_kstride + (size_t)it.z * ds.nb;
    const int cnt0 = flagged_blocks(ftile, Pw + 1, min(ds.nb, Pw + 1 + kQChunk), qlist, &qcount);
    const bool live = r < ds.n;
    double a[QS], cs[NC];
#pragma unroll
    for (int q = 0; q < QS; ++q) a[q] = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) cs[c] = 0.0;
    {
        // row side: the chunk slots the pass wrote (the workgroup's 256 positions share one block)
        const int P = it.z * kRowsPerBlock / R;
        const int nch = (P * R + ds.L - 1) / ds.L;
        const unsigned char *rf = L.colflag + ds.rflag0 + (size_t)k * ds.rflag_kstride + (size_t)P * ds.max_chunks;
        add_row_slots<NR, QS>(rowpart + ds.row0 + (size_t)k * ds.row_kstride, rf, nch, ds.max_chunks,
                              (size_t)ds.nb * R, r, live, a, rlist, &rcount);
    }
    if (live) add_listed_columns<NC>(cbase, R, r, qlist, cnt0, cs);
    add_remaining_columns<NC>(cbase, ftile, R, Pw, ds.nb, r, live, cs, qlist, &qcount);
    if (!live) return;
    const int i = ds.perm[r];
    double b[QS];
#pragma unroll
    for (int q = 0; q < QS; ++q) b[q] = a[q] + cs[q];
    if (GRAD && ds.fast) { // the base-2 pass summed e y (y = r'^2 + dt'^2) and e dt'^2, r' = sx r, dt' = st dt
        const double *pk = params + 6 * (size_t)(ds.param0 + k);
        const double unit = L.b2tab ? 1.0 / kern::kExp2TabSize : 1.0;
        b[1] = from_base2_sq(b[1] - b[2], pk[1] * pk[1], unit);
        b[2] = from_base2_sq(b[2], pk[2] * pk[2], unit);
    }
#pragma unroll
    for (int q = 0; q < QS; ++q) bsum[((size_t)k * QS + q) * ds.npad + i] = b[q];
}

void launch_sym_bsums(const SymLaunch &L, const double *params, bool grad, const double *rowpart,
                      const double *colpart, double *bsum, hipStream_t s) {
    if (grad)
        hipLaunchKernelGGL((sym_bsum_kernel<true>), dim3((unsigned)L.fin_items), dim3(kRowsPerBlock), 0, s, L, params,
                           rowpart, colpart, bsum);
    else
        hipLaunchKernelGGL((sym_bsum_kernel<false>), dim3((unsigned)L.fin_items), dim3(kRowsPerBlock), 0, s, L, params,
                           rowpart, colpart, bsum);
}

} // namespace kern
} // namespace hph
