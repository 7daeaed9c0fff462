// mpf_compare.hip -- difference between two slot states of the same grid,
// on the device: the FP32 solver's phi against the FP64 solver's after the
// same steps, without copying either dense field (201 phases x 67M cells at
// 512x512x256) to the host.
//
// Per cell, the liquid values are compared directly and the two sorted slot
// lists are merged by grain id (a grain stored in one state and not in the
// other compares with 0, the value it has in the reference's dense phi).
// Reductions are in FP64:
//   out[0]  max over cells and phases of |phi_a - phi_b|
//   out[1]  sum over cells and phases of |phi_a - phi_b|
//   out[2]  cells whose sets of stored grains differ
//   out[3]  cells with |phi_a - phi_b| > tol for some phase
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpf {
namespace cmp {

template <typename R>
struct View {
    const R *liq;
    const uint8_t *cnt;
    const uint16_t *sid;
    const R *sval;
    size_t stride;
};

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// out: [0] max |d| as the bits of a non-negative double (ordered like the
// value), [1] sum |d| (double), [2] set differences, [3] cells above tol.
template <typename RA, typename RB>
__global__ __launch_bounds__(256) void k_compare(View<RA> A, View<RB> B, size_t ncell, double tol,
                                                 unsigned long long *out)
{
    double mx = 0.0, sm = 0.0;
    unsigned long long nset = 0, nover = 0;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += nth) {
        double cm = fabs((double)A.liq[c] - (double)B.liq[c]);
        double cs = cm;
        const int na = A.cnt[c], nb = B.cnt[c];
        int ia = 0, ib = 0;
        bool same = na == nb;
        while (ia < na || ib < nb) {
            const uint32_t ga = ia < na ? A.sid[(size_t)ia * A.stride + c] : 0xFFFFFFFFu;
            const uint32_t gb = ib < nb ? B.sid[(size_t)ib * B.stride + c] : 0xFFFFFFFFu;
            double va = 0.0, vb = 0.0;
            if (ga <= gb) va = (double)A.sval[(size_t)ia++ * A.stride + c];
            if (gb <= ga) vb = (double)B.sval[(size_t)ib++ * B.stride + c];
            same = same && ga == gb;
            const double d = fabs(va - vb);
            cm = fmax(cm, d);
            cs += d;
        }
        mx = fmax(mx, cm);
        sm += cs;
        nset += same ? 0ull : 1ull;
        nover += cm > tol ? 1ull : 0ull;
    }
    mx = wave_max(mx);
    sm = wave_sum(sm);
    nset = wave_sum_u(nset);
    nover = wave_sum_u(nover);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(mx));
        atomicAdd((double *)&out[1], sm);
        atomicAdd(&out[2], nset);
        atomicAdd(&out[3], nover);
    }
}

template <typename RA, typename RB>
static hipError_t launch(const void *const *pa, size_t sa, const void *const *pb, size_t sb, size_t ncell, double tol,
                         unsigned long long *out, hipStream_t st)
{
    View<RA> A{(const RA *)pa[0], (const uint8_t *)pa[1], (const uint16_t *)pa[2], (const RA *)pa[3], sa};
    View<RB> B{(const RB *)pb[0], (const uint8_t *)pb[1], (const uint16_t *)pb[2], (const RB *)pb[3], sb};
    const size_t blocks = (ncell + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 8192 ? (blocks ? blocks : 1) : 8192);
    hipLaunchKernelGGL((k_compare<RA, RB>), dim3(grid), dim3(256), 0, st, A, B, ncell, tol, out);
    return hipGetLastError();
}

} // namespace cmp
} // namespace mpf

// pa/pb: liq, cnt, sid, sval of each state; fa/fb: 1 for double values.
// out: 4 zeroed 64-bit words on the device (see the file header).
extern "C" hipError_t mpf_launch_compare(const void *const *pa, int fa, size_t sa, const void *const *pb, int fb,
                                         size_t sb, size_t ncell, double tol, unsigned long long *out,
                                         hipStream_t st)
{
    using namespace mpf::cmp;
    if (fa && fb) return launch<double, double>(pa, sa, pb, sb, ncell, tol, out, st);
    if (fa) return launch<double, float>(pa, sa, pb, sb, ncell, tol, out, st);
    if (fb) return launch<float, double>(pa, sa, pb, sb, ncell, tol, out, st);
    return launch<float, float>(pa, sa, pb, sb, ncell, tol, out, st);
}
