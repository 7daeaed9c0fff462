// mpf_nn.hip -- kernels of the learned surrogate (no counterpart in the
// reference): a slot-equivariant 3D U-Net trained on the solver's own
// trajectories and used to propose 16-step advances.
//
// Tensors are channels-last, [B][X][Y][Z][C] with z fastest among the
// spatial axes; activations and MFMA operands are bf16, accumulation and
// master weights fp32.  Convolutions are implicit GEMMs on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16): M = voxels, N = output channels,
// K = taps x input channels.  Input and output channel counts are multiples
// of 32.  Each axis is padded periodically or with zeros.
//
//   k_conv_fwd     y = act(conv(x, w) + b), 3x3x3 (27 taps) or 1x1x1 (1 tap),
//                  wave tiles of 16-64 voxels x 32-64 channels
//   k_conv_bwd_w   dw[tap][co][ci] += sum_v dy[v][co] x[v + tap][ci] (bf16 outer
//                  products on the matrix cores, fp32 sums)
//   (backward data is k_conv_fwd on the flipped, transposed weights)
//   k_pool/k_up    2x average pooling, 2x nearest upsampling and their adjoints
//   slot kernels   per-cell active grains of a dense phi, their features,
//                  sum pooling, the head's input, targets, applying increments
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>

#include "mpf_nn.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

// LDS transpose read (ds_read_b64_tr_b16): in each group of 16 lanes, lane
// 16g + j receives element j%4 of the 8-byte rows addressed by lanes
// 16g + 4i + j/4, i = 0..3 (value i from the i-th of those lanes).
__device__ __forceinline__ s16x4 lds_tr16(const bf16 *p)
{
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)p);
}

namespace {

__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *(const bf16x8 *)p; }
__device__ __forceinline__ void st8(bf16 *p, bf16x8 v) { *(bf16x8 *)p = v; }

// V consecutive channels (1, or 8 in one 16-byte access) as floats.
template <int V>
struct Chans {
    float v[V];
};

template <int V>
__device__ __forceinline__ Chans<V> ld_ch(const bf16 *p)
{
    Chans<V> r;
    if constexpr (V == 8) {
        const bf16x8 a = ld8(p);
#pragma unroll
        for (int e = 0; e < 8; ++e) r.v[e] = (float)a[e];
    } else {
        r.v[0] = (float)*p;
    }
    return r;
}

template <int V>
__device__ __forceinline__ void st_ch(bf16 *p, const Chans<V> &x)
{
    if constexpr (V == 8) {
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = (bf16)x.v[e];
        st8(p, o);
    } else {
        *p = (bf16)x.v[0];
    }
}

struct Geom {
    int B, X, Y, Z;    // voxel grid
    int cin, cout;     // channels
    int ntap;          // 27 or 1
    int px, py, pz;    // 1: periodic axis, 0: zero padding
};

__device__ __forceinline__ int wrap_axis(int c, int n, int periodic, bool &ok)
{
    if (c >= 0 && c < n) return c;
    if (!periodic) {
        ok = false;
        return 0;
    }
    return c < 0 ? c + n : c - n;
}

// Flat index of the neighbour (b, x+dx, y+dy, z+dz), or -1 outside a
// zero-padded axis.
__device__ __forceinline__ int neighbour(const Geom &g, int b, int x, int y, int z, int dx, int dy, int dz)
{
    bool ok = true;
    const int xx = wrap_axis(x + dx, g.X, g.px, ok);
    const int yy = wrap_axis(y + dy, g.Y, g.py, ok);
    const int zz = wrap_axis(z + dz, g.Z, g.pz, ok);
    return ok ? ((b * g.X + xx) * g.Y + yy) * g.Z + zz : -1;
}

__device__ __forceinline__ void tap_offsets(int ntap, int tap, int &dx, int &dy, int &dz)
{
    if (ntap == 27) {
        dx = tap / 9 - 1;
        dy = (tap / 3) % 3 - 1;
        dz = tap % 3 - 1;
    } else {
        dx = dy = dz = 0;
    }
}

__device__ __forceinline__ void decompose(const Geom &g, int v, int &b, int &x, int &y, int &z)
{
    z = v % g.Z;
    int r = v / g.Z;
    y = r % g.Y;
    r /= g.Y;
    x = r % g.X;
    b = r / g.X;
}

// ------------------------------------------------------------ convolution
// Block: 4 waves; each wave computes (16 * MT) voxels x (16 * NT) output
// channels as MT x NT MFMA tiles of 16 x 16, so a block covers 64 * MT
// voxels.  MFMA operand layout (16x16x32): lane l supplies A row l%16 and B
// column l%16, eight K values each from K group l/16; it receives D rows
// 4*(l/16)+i, column l%16.  The K loop runs over (tap, 32-channel chunk) in
// that order; with PIPE the operands of the next k-step are loaded before the
// MFMAs of the current one (two register sets).  Every variant accumulates
// in the same order, so all give the same bits.
template <int MT, int NT, bool PIPE>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16 *__restrict__ X, const bf16 *__restrict__ W,
                                                  const float *__restrict__ bias, bf16 *__restrict__ Y,
                                                  float *__restrict__ Yf, Geom g, int relu, int nvox)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (blockIdx.x * 4 + wave) * (16 * MT);
    const int n0 = blockIdx.y * (16 * NT);
    const int row = lane & 15, kq = lane >> 4;
    int vb[MT], vx[MT], vy[MT], vz[MT];
    bool vok[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
        const int v = m0 + 16 * t + row;
        vok[t] = v < nvox;
        decompose(g, vok[t] ? v : 0, vb[t], vx[t], vy[t], vz[t]);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u) acc[t][u] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
    const int kchunks = g.cin >> 5;
    const size_t tapw = (size_t)g.cout * g.cin;
    const bf16 *wrow = W + (size_t)(n0 + row) * g.cin + kq * 8;
    int nb[MT]; // neighbour voxel of each row tile at the current tap, -1 in zero padding
    auto set_tap = [&](int tap) {
        int dx, dy, dz;
        tap_offsets(g.ntap, tap, dx, dy, dz);
#pragma unroll
        for (int t = 0; t < MT; ++t) nb[t] = vok[t] ? neighbour(g, vb[t], vx[t], vy[t], vz[t], dx, dy, dz) : -1;
    };
    auto load = [&](int tap, int kc, bf16x8 *a, bf16x8 *b) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
            a[t] = nb[t] >= 0 ? *(const bf16x8 *)(X + (size_t)nb[t] * g.cin + kc * 32 + kq * 8) : zero8;
        const bf16 *wt = wrow + (size_t)tap * tapw + kc * 32;
#pragma unroll
        for (int u = 0; u < NT; ++u) b[u] = *(const bf16x8 *)(wt + (size_t)16 * u * g.cin);
    };
    auto mma = [&](const bf16x8 *a, const bf16x8 *b) {
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int u = 0; u < NT; ++u)
                acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[u], acc[t][u], 0, 0, 0);
    };
    bf16x8 a0[MT], b0[NT];
    if constexpr (!PIPE) {
        for (int tap = 0; tap < g.ntap; ++tap) {
            set_tap(tap);
            for (int kc = 0; kc < kchunks; ++kc) {
                load(tap, kc, a0, b0);
                mma(a0, b0);
            }
        }
    } else {
        // k-steps s (even) and s + 1 alternate between the register sets
        bf16x8 a1[MT], b1[NT];
        const int nsteps = g.ntap * kchunks;
        int tap = 0, kc = 0;
        set_tap(0);
        load(0, 0, a0, b0);
        for (int s = 0; s < nsteps; s += 2) {
            int tap1 = tap, kc1 = kc + 1;
            if (kc1 == kchunks) {
                kc1 = 0;
                ++tap1;
            }
            const bool has1 = s + 1 < nsteps;
            if (has1) {
                if (tap1 != tap) set_tap(tap1);
                load(tap1, kc1, a1, b1);
            }
            mma(a0, b0);
            if (!has1) break;
            int tap2 = tap1, kc2 = kc1 + 1;
            if (kc2 == kchunks) {
                kc2 = 0;
                ++tap2;
            }
            if (s + 2 < nsteps) {
                if (tap2 != tap1) set_tap(tap2);
                load(tap2, kc2, a0, b0);
            }
            mma(a1, b1);
            tap = tap2;
            kc = kc2;
        }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const int c = n0 + 16 * u + row;
            const float bc = bias ? bias[c] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int v = m0 + 16 * t + 4 * kq + i;
                if (v >= nvox) continue;
                float r = acc[t][u][i] + bc;
                if (relu) r = r > 0.f ? r : 0.f;
                if (Y) Y[(size_t)v * g.cout + c] = (bf16)r;
                if (Yf) Yf[(size_t)v * g.cout + c] = r;
            }
        }
}

// dW[tap][co][ci] += sum over a chunk of voxels of dY[v][co] * X[v + tap][ci],
// as sums of outer products on the bf16 matrix cores
// (v_mfma_f32_32x32x16_bf16, K = 16 voxels, fp32 accumulation; the products
// of bf16 values are exact in fp32).  Both tensors are channels-last, so a
// lane's eight K values (eight voxels of one channel) are 64 bytes apart in
// memory: each wave stages its 16 voxels x 32 channels of dY and of X at each
// z-tap in LDS with one 16-byte load per lane and tensor (lane l: voxel l/4,
// channels 8*(l%4)..+7), then lane l reads column l%32, voxels 8*(l/32)..+7
// of the tiles as its A (dY, rows = co) and B (X, columns = ci) operands with
// two LDS transpose reads each (ds_read_b64_tr_b16: 8 reads for NZ = 3 where
// 2-byte reads take 32; 12% faster at 4 x 64^3 voxels, 32 -> 32 channels),
// and receives dW rows 8*(r/4) + 4*(l/32) + r%4 (r < 16) of column l%32.
// Block: 4 waves over a 32 x 32 (co, ci) tile and NZ taps that differ only in
// z (the three z offsets of one (x, y) offset, or the one 1x1x1 tap), which
// share the dY operand and the coordinate arithmetic.  The waves take
// interleaved 16-voxel groups of the block's chunk; their sums are added in
// LDS, then one atomic per weight.  The tiles are per wave and a wave's LDS
// accesses complete in order, so the voxel loop needs only compiler barriers
// (no s_barrier) between the stores and the reads of a tile.
// (A version on the fp32 matrix cores, v_mfma_f32_32x32x2f32 with K = 2
// voxels read straight from memory, ran 2-8.7x slower on the surrogate's
// layer shapes: 0.40 against 0.066 ms at 64^3 voxels, 32 -> 32 channels.)
template <int NZ>
__global__ __launch_bounds__(256) void k_conv_bwd_w(const bf16 *__restrict__ X, const bf16 *__restrict__ dY,
                                                    float *__restrict__ dW, Geom g, int nvox, int vchunk)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 31, kg = lane >> 5, half = kg;
    const int nct = g.cout >> 5;
    const int c0 = (blockIdx.z % nct) * 32, i0 = (blockIdx.z / nct) * 32;
    const int t0 = blockIdx.y * NZ;
    int dx, dy, dz0;
    tap_offsets(g.ntap, t0, dx, dy, dz0);
    __shared__ __align__(16) bf16 s_t[4][1 + NZ][16][32];
    f32x16 acc[NZ];
#pragma unroll
    for (int q = 0; q < NZ; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
    const int vbeg = blockIdx.x * vchunk;
    const int vend = min(vbeg + vchunk, nvox);
    const int vr = lane >> 2, cq = (lane & 3) * 8;
    const bf16x8 zero8 = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
    // Lane l's operand: column l%32, voxels 8*(l/32)..+7 of a 16 x 32 tile,
    // as two transpose reads of 4 voxels.  With g = l/16 and j = l%16, the
    // lanes 16g + 4i + m (i, m < 4) address voxel 8*(g/2) + i (+4 for the
    // second read), channels 16*(g%2) + 4m..+3, so lane 16g + j receives
    // channel 16*(g%2) + j = l%32 of voxels 8*(g/2) + 0..3.
    const int trow = 8 * (lane >> 5) + ((lane & 15) >> 2);
    const int tcol = 16 * ((lane >> 4) & 1) + 4 * (lane & 3);
    auto tile_column = [&](const bf16 (*t)[32]) {
        const s16x4 lo = lds_tr16(&t[trow][tcol]);
        const s16x4 hi = lds_tr16(&t[trow + 4][tcol]);
        const s16x8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
        return __builtin_bit_cast(bf16x8, v);
    };
    for (int v0 = vbeg + 16 * wave; v0 < vend; v0 += 64) {
        const int v = v0 + vr;
        const bool ok = v < vend;
        int b, x, y, z;
        decompose(g, ok ? v : 0, b, x, y, z);
        // element offsets fit in 32 bits (checked by conv_args_ok)
        const bf16x8 ldy = ok ? ld8(dY + (uint32_t)(v * g.cout + c0 + cq)) : zero8;
        bool okxy = ok;
        const int xx = wrap_axis(x + dx, g.X, g.px, okxy);
        const int yy = wrap_axis(y + dy, g.Y, g.py, okxy);
        const int row = ((b * g.X + xx) * g.Y + yy) * g.Z;
        bf16x8 lx[NZ];
#pragma unroll
        for (int q = 0; q < NZ; ++q) {
            bool okz = okxy;
            const int zz = wrap_axis(z + dz0 + q, g.Z, g.pz, okz);
            lx[q] = okz ? ld8(X + (uint32_t)((row + zz) * g.cin + i0 + cq)) : zero8;
        }
        st8(&s_t[wave][0][vr][cq], ldy);
#pragma unroll
        for (int q = 0; q < NZ; ++q) st8(&s_t[wave][1 + q][vr][cq], lx[q]);
        __builtin_amdgcn_wave_barrier();
        bf16x8 a, bq[NZ];
        a = tile_column(s_t[wave][0]);
#pragma unroll
        for (int q = 0; q < NZ; ++q) bq[q] = tile_column(s_t[wave][1 + q]);
#pragma unroll
        for (int q = 0; q < NZ; ++q) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[q], acc[q], 0, 0, 0);
        __builtin_amdgcn_wave_barrier();
    }
    __shared__ float s_red[4][32 * 32];
#pragma unroll
    for (int q = 0; q < NZ; ++q) {
#pragma unroll
        for (int r = 0; r < 16; ++r) s_red[wave][(8 * (r / 4) + 4 * half + (r % 4)) * 32 + col] = acc[q][r];
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const float sum = s_red[0][e] + s_red[1][e] + s_red[2][e] + s_red[3][e];
            atomicAdd(&dW[((size_t)(t0 + q) * g.cout + c0 + (e >> 5)) * g.cin + i0 + (e & 31)], sum);
        }
        __syncthreads();
    }
}

// fp32 master weights [ntap][co][ci] -> bf16, as they are (flip 0) or flipped
// in space and transposed to [ntap][ci][co] (flip 1: the weights of the
// backward-data convolution).
__global__ void k_weights_bf16(const float *W, bf16 *out, int ntap, int co, int ci, int flip)
{
    const size_t n = (size_t)ntap * co * ci;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!flip) {
        out[i] = (bf16)W[i];
        return;
    }
    // out[t][c_in][c_out] = W[ntap-1-t][c_out][c_in]
    const int c_out = (int)(i % co);
    const int c_in = (int)((i / co) % ci);
    const int t = (int)(i / ((size_t)co * ci));
    out[i] = (bf16)W[((size_t)(ntap - 1 - t) * co + c_out) * ci + c_in];
}

__global__ void k_relu_bwd(bf16 *dY, const bf16 *Y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !((float)Y[i] > 0.f)) dY[i] = (bf16)0.f;
}

// db[c] += sum_v dY[v][c]; one block per chunk of voxels per 64 channels.
__global__ void k_bias_grad(const bf16 *dY, float *db, int nvox, int C, int chunk)
{
    const int c = blockIdx.y * 64 + (threadIdx.x & 63);
    const int part = threadIdx.x >> 6;
    __shared__ float red[4][64];
    float s = 0.f;
    if (c < C) {
        const int vbeg = blockIdx.x * chunk;
        for (int v = vbeg + part; v < min(vbeg + chunk, nvox); v += 4) s += (float)dY[(size_t)v * C + c];
    }
    red[part][threadIdx.x & 63] = s;
    __syncthreads();
    if (part == 0 && c < C) atomicAdd(&db[c], red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                                  red[3][threadIdx.x]);
}

// The same sums with 16-byte loads: thread t of a block takes channel group
// t % G (8 channels, G = C / 8 <= 256) of voxel rows t / G, t / G + R, ...
// (R = 256 / G rows per pass), so consecutive threads read consecutive
// 16-byte pieces of dY; the block adds its rows in LDS and issues one atomic
// per channel.  Chunks are sized for a few hundred blocks, which keeps the
// atomics on each channel few.
__global__ __launch_bounds__(256) void k_bias_grad8(const bf16 *dY, float *db, int nvox, int C, int chunk)
{
    const int G = C >> 3, R = 256 / G;
    const int t = threadIdx.x, g = t % G, r = t / G;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (r < R) {
        const int vbeg = blockIdx.x * chunk;
        const int vend = min(vbeg + chunk, nvox);
        for (int v = vbeg + r; v < vend; v += R) {
            const bf16x8 a = ld8(dY + (size_t)v * C + g * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)a[j];
        }
    }
    __shared__ float red[8][257];
#pragma unroll
    for (int j = 0; j < 8; ++j) red[j][t] = s[j];
    __syncthreads();
    for (int c = t; c < C; c += 256) {
        const int gg = c >> 3, j = c & 7;
        float sum = 0.f;
        for (int rr = 0; rr < R; ++rr) sum += red[j][rr * G + gg];
        atomicAdd(&db[c], sum);
    }
}

// 2x average pooling [B][X][Y][Z][C] -> [B][X/2][Y/2][Z/2][C] and its adjoint,
// 2x nearest upsampling [B][x][y][z][C] -> [B][2x][2y][2z][C] and its
// adjoint.  A thread takes V channels of one voxel (V = 8: one 16-byte
// access); I is the index type (32-bit when the host finds it wide enough).
// Per channel the operations are the same for every V.
template <int V, typename I>
__global__ void k_pool_fwd(const bf16 *X, bf16 *Y, int B, int nx, int ny, int nz, int C)
{
    const int hx = nx / 2, hy = ny / 2, hz = nz / 2, CV = C / V;
    const I n = (I)B * hx * hy * hz * CV;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % CV) * V;
    I r = i / CV;
    const int z = (int)(r % hz); r /= hz;
    const int y = (int)(r % hy); r /= hy;
    const int x = (int)(r % hx);
    const int b = (int)(r / hx);
    Chans<V> s;
#pragma unroll
    for (int e = 0; e < V; ++e) s.v[e] = 0.f;
    for (int a = 0; a < 8; ++a) {
        const int xx = 2 * x + (a >> 2), yy = 2 * y + ((a >> 1) & 1), zz = 2 * z + (a & 1);
        const Chans<V> t = ld_ch<V>(X + ((((size_t)b * nx + xx) * ny + yy) * nz + zz) * C + c);
#pragma unroll
        for (int e = 0; e < V; ++e) s.v[e] += t.v[e];
    }
#pragma unroll
    for (int e = 0; e < V; ++e) s.v[e] *= 0.125f;
    st_ch<V>(Y + (size_t)i * V, s);
}

template <int V, typename I>
__global__ void k_pool_bwd(const bf16 *dY, bf16 *dX, int B, int nx, int ny, int nz, int C)
{
    const int CV = C / V;
    const I n = (I)B * nx * ny * nz * CV;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % CV) * V;
    I r = i / CV;
    const int z = (int)(r % nz); r /= nz;
    const int y = (int)(r % ny); r /= ny;
    const int x = (int)(r % nx);
    const int b = (int)(r / nx);
    const int hx = nx / 2, hy = ny / 2, hz = nz / 2;
    Chans<V> t = ld_ch<V>(dY + ((((size_t)b * hx + x / 2) * hy + y / 2) * hz + z / 2) * C + c);
#pragma unroll
    for (int e = 0; e < V; ++e) t.v[e] = 0.125f * t.v[e];
    st_ch<V>(dX + (size_t)i * V, t);
}

template <int V, typename I>
__global__ void k_up_fwd(const bf16 *X, bf16 *Y, int B, int nx, int ny, int nz, int C)
{
    const int fx = 2 * nx, fy = 2 * ny, fz = 2 * nz, CV = C / V;
    const I n = (I)B * fx * fy * fz * CV;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % CV) * V;
    I r = i / CV;
    const int z = (int)(r % fz); r /= fz;
    const int y = (int)(r % fy); r /= fy;
    const int x = (int)(r % fx);
    const int b = (int)(r / fx);
    st_ch<V>(Y + (size_t)i * V, ld_ch<V>(X + ((((size_t)b * nx + x / 2) * ny + y / 2) * nz + z / 2) * C + c));
}

template <int V, typename I>
__global__ void k_up_bwd(const bf16 *dY, bf16 *dX, int B, int nx, int ny, int nz, int C)
{
    const int CV = C / V;
    const I n = (I)B * nx * ny * nz * CV;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % CV) * V;
    I r = i / CV;
    const int z = (int)(r % nz); r /= nz;
    const int y = (int)(r % ny); r /= ny;
    const int x = (int)(r % nx);
    const int b = (int)(r / nx);
    const int fx = 2 * nx, fy = 2 * ny, fz = 2 * nz;
    Chans<V> s;
#pragma unroll
    for (int e = 0; e < V; ++e) s.v[e] = 0.f;
    for (int a = 0; a < 8; ++a) {
        const int xx = 2 * x + (a >> 2), yy = 2 * y + ((a >> 1) & 1), zz = 2 * z + (a & 1);
        const Chans<V> t = ld_ch<V>(dY + ((((size_t)b * fx + xx) * fy + yy) * fz + zz) * C + c);
#pragma unroll
        for (int e = 0; e < V; ++e) s.v[e] += t.v[e];
    }
    st_ch<V>(dX + (size_t)i * V, s);
}

// dst[v][doff + j] (=, or += with acc) src[v][soff + j] for j < nch.
__global__ void k_copy_ch(const bf16 *src, int sC, int soff, bf16 *dst, int dC, int doff, int nch, int nvox, int acc)
{
    const size_t n = (size_t)nvox * nch;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = (int)(i % nch);
    const size_t v = i / nch;
    const float s = (float)src[v * sC + soff + j];
    bf16 *d = dst + v * dC + doff + j;
    *d = acc ? (bf16)((float)*d + s) : (bf16)s;
}


// ------------------------------------------------------------ channel kernels, 8 channels per thread
// Vector forms of the channel kernels below for the common case (channel
// counts and offsets multiples of 8, 16-byte aligned tensors, fewer than 2^31
// vectors): one thread moves 8 bf16 channels (16 bytes) and index arithmetic
// is 32-bit.  Per channel they do the same operations in the same order as
// the scalar forms, so the results are identical.

__global__ void k_copy_ch8(const bf16 *src, int sC, int soff, bf16 *dst, int dC, int doff, int nch8, int n8,
                           int acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int j = (i % nch8) * 8;
    const unsigned v = (unsigned)(i / nch8);
    const bf16x8 a = ld8(src + (size_t)v * sC + soff + j);
    bf16 *d = dst + (size_t)v * dC + doff + j;
    if (!acc) {
        st8(d, a);
        return;
    }
    bf16x8 o = ld8(d);
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (bf16)((float)o[e] + (float)a[e]);
    st8(d, o);
}

__global__ void k_slot_pool8(const bf16 *e, const int32_t *sid, bf16 *h, int nv, int K, int C)
{
    const int G = C / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv * G) return;
    const int v = i / G;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < K; ++k)
        if (sid[(size_t)k * nv + v] >= 0) {
            const bf16x8 a = ld8(e + ((size_t)k * nv * G + i) * 8);
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] += (float)a[c];
        }
    bf16x8 o;
#pragma unroll
    for (int c = 0; c < 8; ++c) o[c] = (bf16)s[c];
    st8(h + (size_t)i * 8, o);
}

__global__ void k_slot_pool_bwd8(const bf16 *dh, const int32_t *sid, bf16 *de, int nv, int K, int C, int acc)
{
    const int G = C / 8;
    const int nvg = nv * G;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvg * K) return;
    const int k = i / nvg, r = i % nvg;
    const int v = r / G;
    const bool on = sid[(size_t)k * nv + v] >= 0;
    const bf16x8 d = ld8(dh + (size_t)r * 8);
    bf16 *p = de + (size_t)i * 8;
    bf16x8 o = acc ? ld8(p) : bf16x8{};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float dv = on ? (float)d[c] : 0.f;
        o[c] = acc ? (bf16)((float)o[c] + dv) : (bf16)dv;
    }
    st8(p, o);
}

__global__ void k_head_in8(const bf16 *F, int CF, const bf16 *e, int CE, bf16 *hin, int nv, int K)
{
    const int G = (CF + CE) / 8, GF = CF / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * nv * G) return;
    const int g = i % G;
    const int sv = i / G;
    const int v = sv % nv;
    st8(hin + (size_t)i * 8, g < GF ? ld8(F + (size_t)v * CF + g * 8) : ld8(e + (size_t)sv * CE + (g - GF) * 8));
}

__global__ void k_head_in_bwd8(const bf16 *dhin, bf16 *dF, int CF, bf16 *de, int CE, int nv, int K)
{
    const int C = CF + CE, G = C / 8, GF = CF / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv * G) return;
    const int g = i % G;
    const int v = i / G;
    if (g < GF) {
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < K; ++k) {
            const bf16x8 a = ld8(dhin + ((size_t)k * nv + v) * C + g * 8);
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] += (float)a[c];
        }
        bf16x8 o;
#pragma unroll
        for (int c = 0; c < 8; ++c) o[c] = (bf16)s[c];
        st8(dF + (size_t)v * CF + g * 8, o);
    } else {
        for (int k = 0; k < K; ++k)
            st8(de + ((size_t)k * nv + v) * CE + (g - GF) * 8, ld8(dhin + ((size_t)k * nv + v) * C + g * 8));
    }
}

__global__ void k_relu_bwd8(bf16 *dY, const bf16 *Y, int n8)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const bf16x8 y = ld8(Y + (size_t)i * 8);
    bf16x8 d = ld8(dY + (size_t)i * 8);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (!((float)y[c] > 0.f)) d[c] = (bf16)0.f;
    st8(dY + (size_t)i * 8, d);
}

// Adam on fp32 parameters (bias-corrected step sizes passed in).
__global__ void k_adam(float *p, const float *gr, float *m, float *v, size_t n, float lr_c, float b1, float b2,
                       float eps, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_c * mi / (sqrtf(vi * c2) + eps);
}

// ------------------------------------------------------------ slot kernels
// The band [z0, z0 + Zb) of a dense phi[N][nx][ny][nz] (float32, the
// reference's layout) as network input.  A cell's slots are the grains
// (id >= 1) with phi > 0 at the cell or at one of its six neighbours (x, y
// periodic, z clamped), ascending id, at most K.  Per slot s:
//   sid[s][v]          grain id or -1
//   feat[s][v][0..31]  phi_g, three orientation features of g, zeros
// and per cell aux[v][0..31] = phi_liquid, temperature feature, 1, zeros.
// *overflow counts cells with more than K candidate grains.
__global__ void k_slot_features(const float *phi, int N, int nx, int ny, int nz, int z0, int Zb, int K,
                                const float *orient, const float *tfeat, int32_t *sid, bf16 *feat, bf16 *aux,
                                int *overflow)
{
    const int nv = nx * ny * Zb;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int kz = v % Zb, m = (v / Zb) % ny, l = v / (Zb * ny);
    const int k = z0 + kz;
    const size_t ncell = (size_t)nx * ny * nz;
    auto at = [&](int g, int ll, int mm, int kk) { return phi[(size_t)g * ncell + ((size_t)ll * ny + mm) * nz + kk]; };
    const int lp = l + 1 < nx ? l + 1 : 0, lm = l > 0 ? l - 1 : nx - 1;
    const int mp = m + 1 < ny ? m + 1 : 0, mmn = m > 0 ? m - 1 : ny - 1;
    const int kp = k + 1 < nz ? k + 1 : nz - 1, km = k > 0 ? k - 1 : 0;
    int cnt = 0;
    bool over = false;
    for (int g = 1; g < N; ++g) {
        const bool act = at(g, l, m, k) > 0.f || at(g, lp, m, k) > 0.f || at(g, lm, m, k) > 0.f ||
                         at(g, l, mp, k) > 0.f || at(g, l, mmn, k) > 0.f || at(g, l, m, kp) > 0.f ||
                         at(g, l, m, km) > 0.f;
        if (!act) continue;
        if (cnt == K) {
            over = true;
            continue;
        }
        sid[(size_t)cnt * nv + v] = g;
        bf16 *f = feat + ((size_t)cnt * nv + v) * 32;
        f[0] = (bf16)at(g, l, m, k);
        f[1] = (bf16)orient[g * 3 + 0];
        f[2] = (bf16)orient[g * 3 + 1];
        f[3] = (bf16)orient[g * 3 + 2];
        for (int j = 4; j < 32; ++j) f[j] = (bf16)0.f;
        ++cnt;
    }
    for (int s = cnt; s < K; ++s) {
        sid[(size_t)s * nv + v] = -1;
        bf16 *f = feat + ((size_t)s * nv + v) * 32;
        for (int j = 0; j < 32; ++j) f[j] = (bf16)0.f;
    }
    if (over) atomicAdd(overflow, 1);
    bf16 *a = aux + (size_t)v * 32;
    a[0] = (bf16)at(0, l, m, k);
    a[1] = (bf16)tfeat[k];
    a[2] = (bf16)1.f;
    for (int j = 3; j < 32; ++j) a[j] = (bf16)0.f;
}

// h[v][c] = sum over valid slots of e[s][v][c]  (e: [K][nv][C]).
__global__ void k_slot_pool(const bf16 *e, const int32_t *sid, bf16 *h, int nv, int K, int C)
{
    const size_t n = (size_t)nv * C;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = (int)(i / C);
    float s = 0.f;
    for (int k = 0; k < K; ++k)
        if (sid[(size_t)k * nv + v] >= 0) s += (float)e[(size_t)k * n + i];
    h[i] = (bf16)s;
}

// de[s][v][c] (+)= dh[v][c] for valid slots, 0 for empty ones.
__global__ void k_slot_pool_bwd(const bf16 *dh, const int32_t *sid, bf16 *de, int nv, int K, int C, int acc)
{
    const size_t n = (size_t)nv * C;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * K) return;
    const int k = (int)(i / n);
    const size_t r = i % n;
    const int v = (int)(r / C);
    const float d = sid[(size_t)k * nv + v] >= 0 ? (float)dh[r] : 0.f;
    de[i] = acc ? (bf16)((float)de[i] + d) : (bf16)d;
}

// Head input hin[s][v] = [F[v][0..CF) , e[s][v][0..CE)] (channels CF + CE).
__global__ void k_head_in(const bf16 *F, int CF, const bf16 *e, int CE, bf16 *hin, int nv, int K)
{
    const int C = CF + CE;
    const size_t n = (size_t)K * nv * C;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const size_t sv = i / C;
    const int v = (int)(sv % nv);
    hin[i] = c < CF ? F[(size_t)v * CF + c] : e[sv * CE + (c - CF)];
}

// Adjoint of k_head_in: dF[v] = sum_s dhin[s][v][0..CF), de[s][v] = dhin[s][v][CF..).
__global__ void k_head_in_bwd(const bf16 *dhin, bf16 *dF, int CF, bf16 *de, int CE, int nv, int K)
{
    const int C = CF + CE;
    const size_t n = (size_t)nv * C;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int v = (int)(i / C);
    if (c < CF) {
        float s = 0.f;
        for (int k = 0; k < K; ++k) s += (float)dhin[((size_t)k * nv + v) * C + c];
        dF[(size_t)v * CF + c] = (bf16)s;
    } else {
        for (int k = 0; k < K; ++k)
            de[((size_t)k * nv + v) * CE + (c - CF)] = dhin[((size_t)k * nv + v) * C + c];
    }
}

// Targets: tgt[s][v] = scale * (phi1[g][cell] - phi0[g][cell]) for valid
// slots (mask 1), 0 with mask 0 otherwise.
__global__ void k_slot_targets(const float *phi0, const float *phi1, int nx, int ny, int nz, int z0, int Zb, int K,
                               const int32_t *sid, float scale, float *tgt, float *mask)
{
    const int nv = nx * ny * Zb;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * nv) return;
    const int v = (int)(i % nv);
    const int g = sid[i];
    if (g < 0) {
        tgt[i] = 0.f;
        mask[i] = 0.f;
        return;
    }
    const int kz = v % Zb, m = (v / Zb) % ny, l = v / (Zb * ny);
    const size_t c = (size_t)g * nx * ny * nz + ((size_t)l * ny + m) * nz + z0 + kz;
    tgt[i] = scale * (phi1[c] - phi0[c]);
    mask[i] = 1.f;
}

// Masked squared error of pred[s][v] = out[(s*nv+v)*stride] against tgt:
// loss partial sums (atomic) and the gradient d/dpred of sum/count written
// into dout[(s*nv+v)*stride] (other channels zero).
__global__ void k_mse(const float *out, int stride, const float *tgt, const float *mask, size_t n, float inv_count,
                      bf16 *dout, float *loss)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float red[256];
    float l = 0.f;
    if (i < n) {
        const float d = (out[i * stride] - tgt[i]) * mask[i];
        l = d * d;
        bf16 *o = dout + i * stride;
        o[0] = (bf16)(2.f * d * inv_count);
        for (int j = 1; j < stride; ++j) o[j] = (bf16)0.f;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Apply predicted increments to the band of phi (in place): each slot's
// grain gets phi + out/scale clipped to [0, 1]; then the liquid is
// 1 - sum of solids, and solids summing above 1 are scaled down to 1.
__global__ void k_slot_apply(float *phi, int N, int nx, int ny, int nz, int z0, int Zb, int K, const int32_t *sid,
                             const float *out, int stride, float inv_scale)
{
    const int nv = nx * ny * Zb;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int kz = v % Zb, m = (v / Zb) % ny, l = v / (Zb * ny);
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t c = ((size_t)l * ny + m) * nz + z0 + kz;
    for (int s = 0; s < K; ++s) {
        const int g = sid[(size_t)s * nv + v];
        if (g < 0) continue;
        float p = phi[(size_t)g * ncell + c] + out[((size_t)s * nv + v) * stride] * inv_scale;
        phi[(size_t)g * ncell + c] = p < 0.f ? 0.f : (p > 1.f ? 1.f : p);
    }
    float sum = 0.f;
    for (int g = 1; g < N; ++g) sum += phi[(size_t)g * ncell + c];
    if (sum > 1.f) {
        const float r = 1.f / sum;
        for (int g = 1; g < N; ++g) phi[(size_t)g * ncell + c] *= r;
        phi[c] = 0.f;
    } else {
        phi[c] = 1.f - sum;
    }
}

// Band reductions over a dense phi[N][nx][ny][nz]: element i of the band
// is column i / Zb (the flat (g, l, m) index) at height z0 + i % Zb, so
// consecutive threads read consecutive heights.  Each thread keeps a running
// maximum over a grid-stride loop (32-bit indices: the host checks the size);
// one block reduction and one atomic per block.
__device__ __forceinline__ void block_max_to(float d, unsigned *out)
{
    __shared__ float red[256];
    red[threadIdx.x] = d;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// max |a - b| over the band of two dense phi, as float bits (non-negative).
__global__ void k_band_maxdiff(const float *a, const float *b, unsigned n, int nz, int z0, int Zb, unsigned *out)
{
    float d = 0.f;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const size_t q = (size_t)(i / Zb) * nz + z0 + i % Zb;
        d = fmaxf(d, fabsf(a[q] - b[q]));
    }
    block_max_to(d, out);
}

// max over the band of |(c - b) - (b - a) * inv_h|: the exact one-step change
// c - b at a proposal b against the proposal's mean rate from a over h steps.
__global__ void k_band_residual(const float *a, const float *b, const float *c, unsigned n, int nz, int z0, int Zb,
                                float inv_h, unsigned *out)
{
    float d = 0.f;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const size_t q = (size_t)(i / Zb) * nz + z0 + i % Zb;
        d = fmaxf(d, fabsf((c[q] - b[q]) - (b[q] - a[q]) * inv_h));
    }
    block_max_to(d, out);
}

// float <-> bf16 conversions of plain arrays.
__global__ void k_to_bf16(const float *x, bf16 *y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

__global__ void k_fill_f32(float *x, float v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// Sum of x[0, n) added to *out (zeroed by the caller): grid-stride partial
// sums, a wave reduction, one atomic per block.  The order of the atomics is
// free, so the result is exact (and the same on every run) for sums of
// integer values below 2^24, such as the counts of a 0/1 mask.
__global__ __launch_bounds__(256) void k_sum_f32(const float *x, size_t n, float *out)
{
    float s = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += x[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, (ws[0] + ws[1]) + (ws[2] + ws[3]));
}

// The 8-channel kernels apply: every pointer 16-byte aligned, every channel
// count and offset a multiple of 8, and fewer than 2^31 vectors.
inline bool vec8(std::initializer_list<const void *> ptrs, std::initializer_list<long long> counts, size_t nvec)
{
    for (const void *p : ptrs)
        if ((uintptr_t)p & 15) return false;
    for (long long c : counts)
        if (c % 8) return false;
    return nvec < (size_t)0x7fffffff;
}

inline Geom geom(const int *dims, int cin, int cout, int ntap, const int *periodic)
{
    Geom g;
    g.B = dims[0];
    g.X = dims[1];
    g.Y = dims[2];
    g.Z = dims[3];
    g.cin = cin;
    g.cout = cout;
    g.ntap = ntap;
    g.px = periodic[0];
    g.py = periodic[1];
    g.pz = periodic[2];
    return g;
}

// Wave tile of the forward convolution: the largest that still gives the
// GPU enough blocks (256 CUs), from timings of the U-Net's layers on MI355X:
// 4x4 tiles pay at 65536 voxels x 64 channels with 64+ input channels, 4x2
// at 2048 blocks of 32 output channels, 2x2 down to 256 blocks, 1x2 below;
// 1x1x1 convolutions (two or three k-steps) run unpipelined.
struct ConvTile {
    int mt, nt;
    bool pipe;
};

inline ConvTile conv_tile(int nvox, int cin, int cout, int ntap)
{
    auto blocks = [&](int mt, int nt) { return (long long)nblk(nvox, 64 * mt) * (cout / (16 * nt)); };
    if (ntap == 1) return {2, 2, false};
    if (cout % 64 == 0 && cin >= 64 && blocks(4, 4) >= 256) return {4, 4, true};
    if (cout % 64 == 0 && blocks(2, 4) >= 256) return {2, 4, true};
    if (cin >= 64 && blocks(4, 2) >= 2048) return {4, 2, true};
    if (blocks(2, 2) >= 256) return {2, 2, true};
    return {1, 2, true};
}

inline int conv_args_ok(const int *dims, int cin, int cout, int ntap)
{
    if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[3] < 1) return 0;
    if (cin < 32 || cout < 32 || (cin & 31) || (cout & 31)) return 0;
    if (ntap != 27 && ntap != 1) return 0;
    const double nv = (double)dims[0] * dims[1] * dims[2] * dims[3];
    return nv * (cin > cout ? cin : cout) < 2147483647.0;
}

} // namespace

extern "C" {

int mpf_nn_conv_fwd(const void *x, const void *w, const float *bias, void *y, float *yf, const int *dims, int cin,
                    int cout, int ntap, const int *periodic, int relu, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !w || (!y && !yf) || !periodic || !conv_args_ok(dims, cin, cout, ntap)) return -1;
    const Geom g = geom(dims, cin, cout, ntap, periodic);
    const int nvox = g.B * g.X * g.Y * g.Z;
    const ConvTile ct = conv_tile(nvox, cin, cout, ntap);
    const dim3 grid(nblk(nvox, 64 * ct.mt), cout / (16 * ct.nt));
#define MPF_CONV(M, N, P)                                                                                     \
    hipLaunchKernelGGL((k_conv_fwd<M, N, P>), grid, dim3(256), 0, st, (const bf16 *)x, (const bf16 *)w, bias, \
                       (bf16 *)y, yf, g, relu, nvox)
    if (!ct.pipe)
        MPF_CONV(2, 2, false);
    else if (ct.mt == 1)
        MPF_CONV(1, 2, true);
    else if (ct.mt == 2 && ct.nt == 2)
        MPF_CONV(2, 2, true);
    else if (ct.mt == 2)
        MPF_CONV(2, 4, true);
    else if (ct.nt == 2)
        MPF_CONV(4, 2, true);
    else
        MPF_CONV(4, 4, true);
#undef MPF_CONV
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_conv_bwd_w(const void *x, const void *dy, float *dw, const int *dims, int cin, int cout, int ntap,
                      const int *periodic, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !dy || !dw || !periodic || !conv_args_ok(dims, cin, cout, ntap)) return -1;
    const Geom g = geom(dims, cin, cout, ntap, periodic);
    const int nvox = g.B * g.X * g.Y * g.Z;
    // voxel chunk per block: about 2048 blocks in all, chunks of 256 to 8192
    const int ngrp = (ntap == 27 ? 9 : 1) * (cout / 32) * (cin / 32);
    long long vc = (long long)nvox * ngrp / 2048;
    vc = vc < 256 ? 256 : (vc > 8192 ? 8192 : vc);
    const int vchunk = (int)(vc / 256 * 256);
    const dim3 grid(nblk(nvox, vchunk), ntap == 27 ? 9 : 1, (cout / 32) * (cin / 32));
    if (ntap == 27)
        hipLaunchKernelGGL((k_conv_bwd_w<3>), grid, dim3(256), 0, st, (const bf16 *)x, (const bf16 *)dy, dw, g, nvox,
                           vchunk);
    else
        hipLaunchKernelGGL((k_conv_bwd_w<1>), grid, dim3(256), 0, st, (const bf16 *)x, (const bf16 *)dy, dw, g, nvox,
                           vchunk);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_weights_bf16(const float *w, void *out, int ntap, int cout, int cin, int flip, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!w || !out || ntap < 1 || cout < 1 || cin < 1) return -1;
    const size_t n = (size_t)ntap * cout * cin;
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(n, 256)), dim3(256), 0, st, w, (bf16 *)out, ntap, cout, cin, flip);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_relu_bwd(void *dy, const void *y, size_t n, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!dy || !y) return -1;
    if (n && vec8({dy, y}, {(long long)(n % 8)}, n / 8))
        hipLaunchKernelGGL(k_relu_bwd8, dim3(nblk(n / 8, 256)), dim3(256), 0, st, (bf16 *)dy, (const bf16 *)y,
                           (int)(n / 8));
    else if (n)
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n, 256)), dim3(256), 0, st, (bf16 *)dy, (const bf16 *)y, n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_bias_grad(const void *dy, float *db, int nvox, int c, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!dy || !db || nvox < 1 || c < 1) return -1;
    if (c <= 256 * 8 && vec8({dy}, {c}, (size_t)nvox * c / 8)) {
        // rows per block pass R = 256 / (c / 8); chunks of whole passes, about
        // 512 blocks in all
        const int R = 256 / (c / 8);
        long long ch = ((long long)nvox + 511) / 512;
        ch = (ch + R - 1) / R * R;
        if (ch < 8LL * R) ch = 8LL * R;
        const int chunk = (int)ch;
        hipLaunchKernelGGL(k_bias_grad8, dim3(nblk(nvox, chunk)), dim3(256), 0, st, (const bf16 *)dy, db, nvox, c,
                           chunk);
    } else {
        const int chunk = 256;
        hipLaunchKernelGGL(k_bias_grad, dim3(nblk(nvox, chunk), (c + 63) / 64), dim3(256), 0, st, (const bf16 *)dy, db,
                           nvox, c, chunk);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_pool(const void *x, void *y, const int *dims, int c, int backward, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !y || !dims || c < 1 || (dims[1] & 1) || (dims[2] & 1) || (dims[3] & 1)) return -1;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2] * dims[3] * c / (backward ? 1 : 8);
    const bool v8 = vec8({x, y}, {c}, n / 8);
#define MPF_POOL(K, V, I)                                                                                       \
    hipLaunchKernelGGL((K<V, I>), dim3(nblk(n / V, 256)), dim3(256), 0, st, (const bf16 *)x, (bf16 *)y, dims[0], \
                       dims[1], dims[2], dims[3], c)
    if (backward && v8)
        MPF_POOL(k_pool_bwd, 8, uint32_t);
    else if (backward)
        MPF_POOL(k_pool_bwd, 1, size_t);
    else if (v8)
        MPF_POOL(k_pool_fwd, 8, uint32_t);
    else
        MPF_POOL(k_pool_fwd, 1, size_t);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_upsample(const void *x, void *y, const int *dims, int c, int backward, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !y || !dims || c < 1) return -1;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2] * dims[3] * c * (backward ? 1 : 8);
    const bool v8 = vec8({x, y}, {c}, n / 8);
    if (backward && v8)
        MPF_POOL(k_up_bwd, 8, uint32_t);
    else if (backward)
        MPF_POOL(k_up_bwd, 1, size_t);
    else if (v8)
        MPF_POOL(k_up_fwd, 8, uint32_t);
    else
        MPF_POOL(k_up_fwd, 1, size_t);
#undef MPF_POOL
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_copy_channels(const void *src, int src_c, int src_off, void *dst, int dst_c, int dst_off, int nch,
                         int nvox, int accumulate, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!src || !dst || nch < 1 || nvox < 1 || src_off + nch > src_c || dst_off + nch > dst_c) return -1;
    const size_t n = (size_t)nvox * nch;
    if (vec8({src, dst}, {src_c, src_off, dst_c, dst_off, nch}, n / 8))
        hipLaunchKernelGGL(k_copy_ch8, dim3(nblk(n / 8, 256)), dim3(256), 0, st, (const bf16 *)src, src_c, src_off,
                           (bf16 *)dst, dst_c, dst_off, nch / 8, (int)(n / 8), accumulate);
    else
        hipLaunchKernelGGL(k_copy_ch, dim3(nblk(n, 256)), dim3(256), 0, st, (const bf16 *)src, src_c, src_off,
                           (bf16 *)dst, dst_c, dst_off, nch, nvox, accumulate);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                int t, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!p || !g || !m || !v || t < 1) return -1;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f / (1.f - powf(b2, (float)t));
    if (n) hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr / c1, b1, b2, eps, c2);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_slot_features(const float *phi, int n_phase, int nx, int ny, int nz, int z0, int zb, int k,
                         const float *orient, const float *tfeat, int32_t *sid, void *feat, void *aux, int *overflow,
                         void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!phi || !orient || !tfeat || !sid || !feat || !aux || !overflow || n_phase < 2 || k < 1 || z0 < 0 ||
        zb < 1 || z0 + zb > nz)
        return -1;
    const int nv = nx * ny * zb;
    hipLaunchKernelGGL(k_slot_features, dim3(nblk(nv, 128)), dim3(128), 0, st, phi, n_phase, nx, ny, nz, z0, zb, k,
                       orient, tfeat, sid, (bf16 *)feat, (bf16 *)aux, overflow);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_slot_pool(const void *e, const int32_t *sid, void *h, int nv, int k, int c, int backward, int accumulate,
                     void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!e || !sid || !h || nv < 1 || k < 1 || c < 1) return -1;
    const bool v8 = vec8({e, h}, {c}, (size_t)nv * c * k / 8);
    if (backward && v8)   // e is dh, h is de
        hipLaunchKernelGGL(k_slot_pool_bwd8, dim3(nblk((size_t)nv * c * k / 8, 256)), dim3(256), 0, st,
                           (const bf16 *)e, sid, (bf16 *)h, nv, k, c, accumulate);
    else if (backward)
        hipLaunchKernelGGL(k_slot_pool_bwd, dim3(nblk((size_t)nv * c * k, 256)), dim3(256), 0, st,
                           (const bf16 *)e, sid, (bf16 *)h, nv, k, c, accumulate);
    else if (v8)
        hipLaunchKernelGGL(k_slot_pool8, dim3(nblk((size_t)nv * c / 8, 256)), dim3(256), 0, st, (const bf16 *)e,
                           sid, (bf16 *)h, nv, k, c);
    else
        hipLaunchKernelGGL(k_slot_pool, dim3(nblk((size_t)nv * c, 256)), dim3(256), 0, st, (const bf16 *)e, sid,
                           (bf16 *)h, nv, k, c);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_head_in(const void *f, int cf, const void *e, int ce, void *hin, int nv, int k, int backward,
                   void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!f || !e || !hin || nv < 1 || k < 1) return -1;
    const bool v8 = vec8({f, e, hin}, {cf, ce}, (size_t)k * nv * (cf + ce) / 8);
    if (backward && v8)   // hin is dhin, f is dF, e is de
        hipLaunchKernelGGL(k_head_in_bwd8, dim3(nblk((size_t)nv * (cf + ce) / 8, 256)), dim3(256), 0, st,
                           (const bf16 *)hin, (bf16 *)f, cf, (bf16 *)e, ce, nv, k);
    else if (backward)
        hipLaunchKernelGGL(k_head_in_bwd, dim3(nblk((size_t)nv * (cf + ce), 256)), dim3(256), 0, st,
                           (const bf16 *)hin, (bf16 *)f, cf, (bf16 *)e, ce, nv, k);
    else if (v8)
        hipLaunchKernelGGL(k_head_in8, dim3(nblk((size_t)k * nv * (cf + ce) / 8, 256)), dim3(256), 0, st,
                           (const bf16 *)f, cf, (const bf16 *)e, ce, (bf16 *)hin, nv, k);
    else
        hipLaunchKernelGGL(k_head_in, dim3(nblk((size_t)k * nv * (cf + ce), 256)), dim3(256), 0, st,
                           (const bf16 *)f, cf, (const bf16 *)e, ce, (bf16 *)hin, nv, k);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_slot_targets(const float *phi0, const float *phi1, int nx, int ny, int nz, int z0, int zb, int k,
                        const int32_t *sid, float scale, float *tgt, float *mask, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!phi0 || !phi1 || !sid || !tgt || !mask || z0 < 0 || zb < 1 || z0 + zb > nz) return -1;
    const size_t n = (size_t)k * nx * ny * zb;
    hipLaunchKernelGGL(k_slot_targets, dim3(nblk(n, 256)), dim3(256), 0, st, phi0, phi1, nx, ny, nz, z0, zb, k, sid,
                       scale, tgt, mask);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_mse(const float *out, int stride, const float *tgt, const float *mask, size_t n, float inv_count,
               void *dout, float *loss, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!out || !tgt || !mask || !dout || !loss || stride < 1) return -1;
    if (n) hipLaunchKernelGGL(k_mse, dim3(nblk(n, 256)), dim3(256), 0, st, out, stride, tgt, mask, n, inv_count,
                              (bf16 *)dout, loss);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_slot_apply(float *phi, int n_phase, int nx, int ny, int nz, int z0, int zb, int k, const int32_t *sid,
                      const float *out, int stride, float inv_scale, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!phi || !sid || !out || z0 < 0 || zb < 1 || z0 + zb > nz || stride < 1) return -1;
    const int nv = nx * ny * zb;
    hipLaunchKernelGGL(k_slot_apply, dim3(nblk(nv, 256)), dim3(256), 0, st, phi, n_phase, nx, ny, nz, z0, zb, k,
                       sid, out, stride, inv_scale);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_band_maxdiff(const float *a, const float *b, int n_phase, int nx, int ny, int nz, int z0, int zb,
                        unsigned *out, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!a || !b || !out || z0 < 0 || zb < 1 || z0 + zb > nz) return -1;
    const size_t n = (size_t)n_phase * nx * ny * zb;
    if (n >= 0xffffffffull - 256ull * 2048) return -1;
    const unsigned nb = nblk(n, 256) < 2048 ? nblk(n, 256) : 2048;
    hipLaunchKernelGGL(k_band_maxdiff, dim3(nb), dim3(256), 0, st, a, b, (unsigned)n, nz, z0, zb, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_band_residual(const float *a, const float *b, const float *c, int n_phase, int nx, int ny, int nz,
                         int z0, int zb, float inv_h, unsigned *out, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!a || !b || !c || !out || z0 < 0 || zb < 1 || z0 + zb > nz) return -1;
    const size_t n = (size_t)n_phase * nx * ny * zb;
    if (n >= 0xffffffffull - 256ull * 2048) return -1;
    const unsigned nb = nblk(n, 256) < 2048 ? nblk(n, 256) : 2048;
    hipLaunchKernelGGL(k_band_residual, dim3(nb), dim3(256), 0, st, a, b, c, (unsigned)n, nz, z0, zb, inv_h, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_to_bf16(const float *x, void *y, size_t n, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !y) return -1;
    if (n) hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n, 256)), dim3(256), 0, st, x, (bf16 *)y, n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_fill(float *x, float v, size_t n, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x) return -1;
    if (n) hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n, 256)), dim3(256), 0, st, x, v, n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int mpf_nn_sum(const float *x, size_t n, float *out, void *stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!x || !out) return -1;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st, out, 0.f, (size_t)1);
    if (n) hipLaunchKernelGGL(k_sum_f32, dim3(nblk(n, 256) < 1024 ? nblk(n, 256) : 1024), dim3(256), 0, st, x, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

} // extern "C"
