// Learned correction of the ILU application: a 3-layer message-passing network over
// the matrix graph (width 32, BF16 weights on the matrix cores with FP32 accumulation),
// its backward pass and an Adam training loop, all on the device.
//
// Per application to a residual r whose ILU output is z0 (A with diagonal d):
//   s = |d|^-1/2, Ahat = S A S;  x = [s r, z0 / s] / sigma,  sigma = rms(s r)
//   h1 = relu(x W1s + (Ahat x) W1n + b1)                    (2 -> 32, FP32 FMAs)
//   h2 = relu(h1 W2s + (Ahat h1) W2n + b2)                  (64 -> 32, matrix cores)
//   h3 = relu(h2 W3s + (Ahat h2) W3n + b3)
//   z  = z0 + sigma * s * (h3 wo + bo)
// The network is homogeneous (scaling r scales the correction) and with wo = bo = 0 it
// is plain ILU.  Activations are stored in BF16 (64 bytes per node), the aggregated
// inputs of the matrix-core layers are rounded to BF16 as their operands, and
// accumulation is FP32.
//
// The layers take the nodes in tiles of 32 consecutive matrix rows, and for inference
// the tiles in a breadth-first (Cuthill-McKee) order of the graph between tiles: the
// records of a node's neighbours then lie near its own and are gathered from the L2 of
// the XCD that handles that band of tiles (a color ordering puts every neighbour in
// another color, far away), while the inputs and the correction still move as whole
// 128-byte lines of the matrix-ordered vectors.  Training keeps the matrix order.  A
// tile's graph is stored slot-major (entry k of its 32 nodes side by side), each entry
// the neighbour's position in the layers' order and the weight a_ij s_j; tile widths
// are padded to multiples of 8 with entries (own position, 0) so that every load of the
// aggregation is unconditional.
//
// Parameter vector (FP32, kParams values): W1s[2][32] W1n[2][32] b1[32] W2s[32][32]
// W2n[32][32] b2[32] W3s[32][32] W3n[32][32] b3[32] wo[32] bo.
#include "dev_common.hpp"
#include "solvers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace senk_amd {

namespace gnn {
constexpr int W = 32;  // width
constexpr int oW1s = 0, oW1n = 64, ob1 = 128, oW2s = 160, oW2n = 1184, ob2 = 2208, oW3s = 2240, oW3n = 3264,
              ob3 = 4288, owo = 4320, obo = 4352;
constexpr int kParams = 4353;
constexpr int kSlot = 7;          // partial-sum slot of the context used by the network
constexpr int kGradBlocks = 256;  // workgroups of the weight-gradient reductions
}  // namespace gnn

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// BF16 with round-to-nearest-even: the conversion instruction of gfx950
// (v_cvt_pk_bf16_f32), two values per instruction in bf16x2.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ __forceinline__ unsigned bf16x2(float lo, float hi) {
    const bf16x2_t v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
// Stored position of feature f (= 8t + 4h + i) inside a node's 32 BF16 features:
// 16h + 4t + i, so that the 16 features one lane feeds to the matrix core (half h,
// t = 0..3) are 32 contiguous bytes.
__device__ __forceinline__ int fpos(int f) { return ((f >> 2) & 1) * 16 + (f >> 3) * 4 + (f & 3); }

// ----------------------------------------------------------------- state
// one entry of the layers' graph: neighbour position in the layers' node order and the
// weight a_ij s_j
struct __align__(8) GnnEdge {
    int col;
    float w;
};
struct GnnGraph {
    int N = 0;
    const int *rptr = nullptr, *cind = nullptr;    // A in matrix order (training's products)
    const float *val = nullptr;
    const int *trptr = nullptr, *tcind = nullptr;  // A^T (== A's arrays when A is symmetric)
    const float *tval = nullptr;
    float *s = nullptr;                            // owned; |d|^-1/2 in the layers' order
    int *own_trptr = nullptr, *own_tcind = nullptr;
    float *own_tval = nullptr;
    int *tperm = nullptr;     // owned; layers' tile T = matrix rows 32 tperm[T] + [0, 32) (null: T)
    int *tptr = nullptr;      // owned; first slot of each tile of 32 nodes (ntiles + 1, multiples of 8)
    GnnEdge *ent = nullptr;   // owned; entry (tptr[t] + k) * 32 + l: k-th entry of node 32t + l
};
struct GnnWork {
    int N = 0;
    bool train = false;
    float2 *xin = nullptr;                         // [s r, z0/s] in the layers' order
    unsigned short *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *m2 = nullptr, *m3 = nullptr;
    float *chat = nullptr;
    float2 *x = nullptr, *m1 = nullptr;            // training: normalised inputs and Ahat x
    float *G = nullptr, *G2 = nullptr, *P = nullptr, *Q = nullptr, *dchat = nullptr;
    float *z = nullptr, *e = nullptr, *t = nullptr, *tmp = nullptr;
    double *sc = nullptr;                          // [0] sigma [1] 1/sigma [2] rr [3] ee [4] loss sum
    float *gpart = nullptr;                        // kGradBlocks x kParams
};

}  // namespace senk_amd

struct senk_gnn_s {
    senk_amd::Ctx *c = nullptr;
    float *theta = nullptr;     // FP32 master parameters
    float *theta_r = nullptr;   // the same with the weights rounded to BF16 (used by the kernels)
    senk_amd::bf16x4 *wb = nullptr;  // packed matrix-core operands of layers 2 and 3 (2 x 512)
    double *grad = nullptr;     // accumulated gradient (kParams)
    float *adam_m = nullptr, *adam_v = nullptr;
    long long adam_t = 0;
    senk_amd::GnnGraph g;       // graph bound for inference
    senk_amd::GnnWork w;
    senk_amd::LearnedPrec lp;   // lp.net == this
};

namespace senk_amd {

namespace {
using namespace gnn;

// ----------------------------------------------------------------- parameter packing
// theta_r = theta with the weights rounded to BF16; wb = matrix-core B operands:
// wb[layer][t*64 + lane] = W[8t + 4*(lane/32) + i][lane%32], i < 4, with W = [Ws; Wn].
__global__ void k_pack_params(const float *__restrict__ theta, float *theta_r, bf16x4 *wb) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < kParams) {
        const bool bias = (tid >= ob1 && tid < oW2s) || (tid >= ob2 && tid < oW3s) || (tid >= ob3 && tid < owo) ||
                          tid == obo;
        theta_r[tid] = bias ? theta[tid] : bf2f(f2bf(theta[tid]));
    }
    if (tid < 2 * 512) {
        const int layer = tid / 512, q = tid % 512, t = q / 64, lane = q % 64;
        const int base = layer == 0 ? oW2s : oW3s;  // Ws then Wn follow each other (32x32 each)
        bf16x4 v;
        for (int i = 0; i < 4; i++) {
            const int k = 8 * t + 4 * (lane / 32) + i, o = lane % 32;
            v[i] = (short)f2bf(theta[base + k * W + o]);
        }
        wb[tid] = v;
    }
}

// ----------------------------------------------------------------- forward
// matrix row of the layers' node u
__device__ __forceinline__ long long gnn_row(const int *__restrict__ tperm, long long u) {
    return tperm ? (long long)tperm[u >> 5] * 32 + (u & 31) : u;
}
// xin[u] = [s r, z0/s] at the matrix row of u; partials of (s r)^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gnn_in(int N, const T *__restrict__ r, const double *in_scale,
                                                   const T *__restrict__ z0, const float *__restrict__ s,
                                                   const int *__restrict__ tperm, float2 *xin, double *partials) {
    const double sc = in_scale ? *in_scale : 1.0;
    double ss = 0;
    for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long long)gridDim.x * kBlock) {
        const long long i = gnn_row(tperm, u);
        const float si = s[u];
        const float rh = si * (float)((double)r[i] * sc), zh = (float)z0[i] / si;
        xin[u] = make_float2(rh, zh);
        ss += (double)rh * (double)rh;
    }
    const double v[1] = {ss};
    store_partials<1>(v, partials);
}
__global__ __launch_bounds__(kBlock) void k_gnn_sigma(const double *P, int B, int N, double *sc) {
    const double ss = sum_partials(P, B);
    if (threadIdx.x == 0) {
        const double sg = sqrt(ss / N);
        sc[0] = sg;
        sc[1] = sg > 0 ? 1.0 / sg : 0.0;
    }
}
// Orders this wave's LDS writes before its later LDS reads of other lanes' data (the
// buffers below are per wave, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// layer 1, one thread per node (two tiles per wave); a wave's 64 node records go out
// through the LDS so that consecutive lanes store consecutive 16 bytes
__global__ __launch_bounds__(kBlock) void k_gnn_l1(int N, const int *__restrict__ tptr,
                                                   const GnnEdge *__restrict__ ent, const float *__restrict__ s,
                                                   const float2 *__restrict__ xin, const double *sc,
                                                   const float *__restrict__ th, unsigned short *h1, float2 *xkeep,
                                                   float2 *m1keep) {
    __shared__ __align__(16) uint4 stage[kWavesPerBlock][64 * 4];
    const int u0 = xcd_block() * kBlock + (int)(threadIdx.x & ~63);
    const int u = u0 + (int)(threadIdx.x & 63);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (u0 >= N) return;
    const int t = min(u, N - 1) >> 5;
    const float inv = (float)sc[1];
    const int p0 = tptr[t], w = tptr[t + 1] - p0;
    const GnnEdge *e = ent + (size_t)p0 * 32 + (u & 31);
    float a0 = 0, a1 = 0;
    // widths are multiples of 8: the entries of 8 neighbours are requested at once, then
    // their inputs
    for (int k0 = 0; k0 < w; k0 += 8) {
        GnnEdge ev[8];
        float2 xv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) ev[q] = e[(size_t)(k0 + q) * 32];
#pragma unroll
        for (int q = 0; q < 8; q++) xv[q] = xin[ev[q].col];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            a0 += ev[q].w * xv[q].x;
            a1 += ev[q].w * xv[q].y;
        }
    }
    const int uc = min(u, N - 1);
    const float f = s[uc] * inv;
    const float m0 = a0 * f, m1 = a1 * f;
    const float2 xv = xin[uc];
    const float x0 = xv.x * inv, x1 = xv.y * inv;
    if (xkeep && u < N) {
        xkeep[u] = make_float2(x0, x1);
        m1keep[u] = make_float2(m0, m1);
    }
    float out[W];  // in stored order
#pragma unroll
    for (int o = 0; o < W; o++) {
        float v = th[ob1 + o];
        v += x0 * th[oW1s + o] + x1 * th[oW1s + W + o];
        v += m0 * th[oW1n + o] + m1 * th[oW1n + W + o];
        out[((o >> 2) & 1) * 16 + (o >> 3) * 4 + (o & 3)] = fmaxf(v, 0.f);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint4 o4;
        o4.x = bf16x2(out[8 * q + 0], out[8 * q + 1]);
        o4.y = bf16x2(out[8 * q + 2], out[8 * q + 3]);
        o4.z = bf16x2(out[8 * q + 4], out[8 * q + 5]);
        o4.w = bf16x2(out[8 * q + 6], out[8 * q + 7]);
        stage[wave][lane * 4 + q] = o4;
    }
    wave_lds_sync();
    uint4 *dst = reinterpret_cast<uint4 *>(h1 + (size_t)u0 * W);
    const int nq = min(64, N - u0) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k * 64 + lane < nq) dst[k * 64 + lane] = stage[wave][k * 64 + lane];
}
// four BF16 values from two packed words
__device__ __forceinline__ bf16x4 bf16x4_of(uint2 v) {
    return bf16x4{(short)(v.x & 0xFFFF), (short)(v.x >> 16), (short)(v.y & 0xFFFF), (short)(v.y >> 16)};
}
// Layers 2 and 3 on the matrix cores: one wave per tile of 32 nodes.  Lane l feeds node
// n0 + l%32; its A operand for k-step t (t < 4: own features, t >= 4: aggregated) is
// features 8t + 4*(l/32) + [0,4) of that node (stored contiguously, see fpos).  D element
// (row, col) sits in lane col + 32*((row/4)%2), register (row%4) + 4*(row/8).  Lanes of
// the last tile past N read node N-1 and the tile's padding entries and store nothing.
// The aggregation runs in another lane layout (lanes 2k, 2k+1: node k) and its sums are
// permuted across lanes into this one: two lanes reading the halves of one record touch
// one cache line per instruction where the matrix-core layout (lanes k, k+32) touches it
// twice; measured, 159 against 170 us for layer 2 at 128^3.
template <bool LAST>
__global__ __launch_bounds__(kBlock) void k_gnn_mc(int N, const int *__restrict__ tptr,
                                                   const GnnEdge *__restrict__ ent, const float *__restrict__ s,
                                                   const unsigned short *__restrict__ hin, const bf16x4 *__restrict__ wb,
                                                   const float *__restrict__ bias, unsigned short *hout,
                                                   unsigned short *mkeep, const float *__restrict__ wo,
                                                   const float *__restrict__ bo, float *chat) {
    __shared__ float red[LAST ? kWavesPerBlock : 1][32][33];
    __shared__ __align__(16) unsigned short tileh[kWavesPerBlock][32 * W];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntiles = (N + 31) / 32;
    const int tl = __builtin_amdgcn_readfirstlane(xcd_block() * kWavesPerBlock + wave);
    if (tl >= ntiles) return;
    const int n0 = tl * 32;
    const int half = lane >> 5, col = lane & 31;
    const int node = n0 + col;
    const int nodec = min(node, N - 1);
    // matrix-core B operands (the same for every tile) before the aggregation's loads
    bf16x4 bop[8];
#pragma unroll
    for (int t = 0; t < 8; t++) bop[t] = wb[t * 64 + lane];
    const float b = bias[col];
    bf16x4 a[8];
    float m[16];
#pragma unroll
    for (int q = 0; q < 16; q++) m[q] = 0;
    // own features: 32 contiguous bytes (stored order), k-steps 0..3
    {
        const uint4 *hs = reinterpret_cast<const uint4 *>(hin + (size_t)nodec * W + 16 * half);
        const uint4 u0 = hs[0], u1 = hs[1];
        a[0] = bf16x4_of(make_uint2(u0.x, u0.y));
        a[1] = bf16x4_of(make_uint2(u0.z, u0.w));
        a[2] = bf16x4_of(make_uint2(u1.x, u1.y));
        a[3] = bf16x4_of(make_uint2(u1.z, u1.w));
    }
    // aggregation in the gather layout: lanes 2k and 2k+1 take the two halves of node k's
    // neighbour records (one cache line per record and instruction), then the sums move to
    // the matrix-core layout (lane l: node l%32, half l/32)
    const int gnode = lane >> 1, ghalf = lane & 1;
    const int p0 = tptr[tl], w = tptr[tl + 1] - p0;
    const GnnEdge *e = ent + (size_t)p0 * 32 + gnode;
    for (int k8 = 0; k8 < w; k8 += 8) {
        GnnEdge ev[8];
#pragma unroll
        for (int q = 0; q < 8; q++) ev[q] = e[(size_t)(k8 + q) * 32];
#pragma unroll
        for (int q4 = 0; q4 < 8; q4 += 4) {
            uint4 hv[4][2];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 *hj = reinterpret_cast<const uint4 *>(hin + (size_t)ev[q4 + q].col * W + 16 * ghalf);
                hv[q][0] = hj[0];
                hv[q][1] = hj[1];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float wgt = ev[q4 + q].w;
                const unsigned wd[8] = {hv[q][0].x, hv[q][0].y, hv[q][0].z, hv[q][0].w,
                                        hv[q][1].x, hv[q][1].y, hv[q][1].z, hv[q][1].w};
#pragma unroll
                for (int v = 0; v < 8; v++) {
                    m[2 * v] += wgt * bf2f((unsigned short)(wd[v] & 0xFFFF));
                    m[2 * v + 1] += wgt * bf2f((unsigned short)(wd[v] >> 16));
                }
            }
        }
    }
    {
        const int src = (2 * col + half) * 4;
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(m[q])));
    }
    const float si = s[nodec];
#pragma unroll
    for (int q = 0; q < 16; q++) m[q] *= si;
    const bool act = node < N;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const bf16x4 v = bf16x4_of(make_uint2(bf16x2(m[4 * t], m[4 * t + 1]), bf16x2(m[4 * t + 2], m[4 * t + 3])));
        a[4 + t] = v;
        if (mkeep && act) *reinterpret_cast<bf16x4 *>(mkeep + (size_t)node * W + 16 * half + 4 * t) = v;
    }
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[t], bop[t], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = 8 * (r >> 2) + 4 * half + (r & 3);
        const unsigned short hv = f2bf(fmaxf(acc[r] + b, 0.f));
        if (hout) tileh[wave][row * W + fpos(col)] = hv;
        if (LAST) red[wave][row][col] = bf2f(hv) * wo[col];
    }
    wave_lds_sync();
    if (hout) {
        // the tile's 32 node records (64 bytes each, stored order) leave the LDS as two
        // 16-byte stores per lane, consecutive lanes writing consecutive bytes
        const int row = lane >> 1, hh = lane & 1;
        if (n0 + row < N) {
            const uint4 *src = reinterpret_cast<const uint4 *>(&tileh[wave][row * W + 16 * hh]);
            uint4 *dst = reinterpret_cast<uint4 *>(hout + (size_t)(n0 + row) * W + 16 * hh);
            dst[0] = src[0];
            dst[1] = src[1];
        }
    }
    if (LAST) {
        if (lane < 32 && n0 + lane < N) {
            float sum = 0;
#pragma unroll 8
            for (int q = 0; q < 32; q++) sum += red[wave][lane][q];
            chat[n0 + lane] = sum + bo[0];
        }
    }
}
// z at the matrix row of u += sigma * s[u] * chat[u]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gnn_out(int N, const float *__restrict__ s, const int *__restrict__ tperm,
                                                    const float *__restrict__ chat, const double *sc, T *z) {
    const float sg = (float)sc[0];
    for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < N; u += (long long)gridDim.x * kBlock) {
        const long long i = gnn_row(tperm, u);
        z[i] = (T)((double)z[i] + (double)(sg * s[u] * chat[u]));
    }
}

// ----------------------------------------------------------------- backward
// dchat = dz * sigma * s;  G = (dchat wo) * [h3 > 0]
__global__ __launch_bounds__(kBlock) void k_gb_out(int N, const float *__restrict__ dz, const float *__restrict__ s,
                                                   const double *sc, const float *__restrict__ th,
                                                   const unsigned short *__restrict__ h3, float *dchat, float *G) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const float d = dz[i] * (float)sc[0] * s[i];
    dchat[i] = d;
#pragma unroll
    for (int o = 0; o < W; o++) G[(size_t)i * W + o] = bf2f(h3[(size_t)i * W + fpos(o)]) > 0 ? d * th[owo + o] : 0.f;
}
// P = G Ws^T, Q = G Wn^T (per node)
__global__ __launch_bounds__(kBlock) void k_gb_pq(int N, const float *__restrict__ G, const float *__restrict__ Ws,
                                                  const float *__restrict__ Wn, float *P, float *Q) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    float g[W];
#pragma unroll
    for (int o = 0; o < W; o++) g[o] = G[(size_t)i * W + o];
    for (int k = 0; k < W; k++) {
        float p = 0, q = 0;
#pragma unroll
        for (int o = 0; o < W; o++) {
            p += g[o] * Ws[k * W + o];
            q += g[o] * Wn[k * W + o];
        }
        P[(size_t)i * W + k] = p;
        Q[(size_t)i * W + k] = q;
    }
}
// Gprev = (P + Ahat^T Q) * [hprev > 0]
__global__ __launch_bounds__(kBlock) void k_gb_agg(int N, const int *__restrict__ trptr, const int *__restrict__ tcind,
                                                   const float *__restrict__ tval, const float *__restrict__ s,
                                                   const float *__restrict__ P, const float *__restrict__ Q,
                                                   const unsigned short *__restrict__ hprev, float *Gprev) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    float acc[W];
#pragma unroll
    for (int k = 0; k < W; k++) acc[k] = 0;
    for (int j = trptr[i]; j < trptr[i + 1]; j++) {
        const int cj = tcind[j];
        const float wgt = tval[j] * s[cj];
#pragma unroll
        for (int k = 0; k < W; k++) acc[k] += wgt * Q[(size_t)cj * W + k];
    }
    const float si = s[i];
#pragma unroll
    for (int k = 0; k < W; k++) {
        const float dh = P[(size_t)i * W + k] + si * acc[k];
        Gprev[(size_t)i * W + k] = bf2f(hprev[(size_t)i * W + fpos(k)]) > 0 ? dh : 0.f;
    }
}
// Weight gradients of a matrix-core layer over a contiguous node range per workgroup:
// dWs[k][o] = sum hin[k] G[o], dWn[k][o] = sum m[k] G[o], db[o] = sum G[o].
__global__ __launch_bounds__(kBlock) void k_gb_wmc(int N, const unsigned short *__restrict__ hin,
                                                   const unsigned short *__restrict__ mag, const float *__restrict__ G,
                                                   int oWs, int oWn, int ob, float *gpart) {
    __shared__ float th[32][W + 1], tm[32][W + 1], tg[32][W + 1];
    const int per = (N + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * per, hi = min(N, lo + per);
    const int k = threadIdx.x / 8, o0 = (threadIdx.x % 8) * 4;
    float as[4] = {0, 0, 0, 0}, an[4] = {0, 0, 0, 0}, ab = 0;
    for (int n0 = lo; n0 < hi; n0 += 32) {
        for (int q = threadIdx.x; q < 32 * W; q += kBlock) {
            const int r = q / W, f = q % W, n = n0 + r;
            const bool ok = n < hi;
            th[r][f] = ok ? bf2f(hin[(size_t)n * W + fpos(f)]) : 0.f;
            tm[r][f] = ok ? bf2f(mag[(size_t)n * W + fpos(f)]) : 0.f;
            tg[r][f] = ok ? G[(size_t)n * W + f] : 0.f;
        }
        __syncthreads();
        for (int r = 0; r < 32; r++) {
            const float hk = th[r][k], mk = tm[r][k];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                as[q] += hk * tg[r][o0 + q];
                an[q] += mk * tg[r][o0 + q];
            }
            if (threadIdx.x < W) ab += tg[r][threadIdx.x];
        }
        __syncthreads();
    }
    float *gp = gpart + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        gp[oWs + k * W + o0 + q] = as[q];
        gp[oWn + k * W + o0 + q] = an[q];
    }
    if (threadIdx.x < W) gp[ob + threadIdx.x] = ab;
}
// Layer-1 and output-layer gradients: dW1s[c][o] = sum x[c] G1[o], dW1n[c][o] = sum m1[c] G1[o],
// db1 = sum G1, dwo[o] = sum h3[o] dchat, dbo = sum dchat.
__global__ __launch_bounds__(kBlock) void k_gb_w1o(int N, const float2 *__restrict__ x, const float2 *__restrict__ m1,
                                                   const float *__restrict__ G1, const unsigned short *__restrict__ h3,
                                                   const float *__restrict__ dchat, float *gpart) {
    // thread t < 32: output column t of layer 1; t in [32, 64): wo[t-32]; t == 64: bo
    const int per = (N + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * per, hi = min(N, lo + per);
    const int t = threadIdx.x;
    float a[5] = {0, 0, 0, 0, 0};
    if (t < W) {
        for (int n = lo; n < hi; n++) {
            const float g = G1[(size_t)n * W + t];
            const float2 xv = x[n], mv = m1[n];
            a[0] += xv.x * g;
            a[1] += xv.y * g;
            a[2] += mv.x * g;
            a[3] += mv.y * g;
            a[4] += g;
        }
    } else if (t < 2 * W) {
        for (int n = lo; n < hi; n++) a[0] += bf2f(h3[(size_t)n * W + fpos(t - W)]) * dchat[n];
    } else if (t == 2 * W) {
        for (int n = lo; n < hi; n++) a[0] += dchat[n];
    }
    float *gp = gpart + (size_t)blockIdx.x * kParams;
    if (t < W) {
        gp[oW1s + t] = a[0];
        gp[oW1s + W + t] = a[1];
        gp[oW1n + t] = a[2];
        gp[oW1n + W + t] = a[3];
        gp[ob1 + t] = a[4];
    } else if (t < 2 * W) {
        gp[owo + t - W] = a[0];
    } else if (t == 2 * W) {
        gp[obo] = a[0];
    }
}
// grad[p] += sum over workgroups of gpart[wg][p] (fixed order)
__global__ __launch_bounds__(kBlock) void k_gb_reduce(const float *__restrict__ gpart, int nb, double *grad) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kParams) return;
    double s = 0;
    for (int b = 0; b < nb; b++) s += gpart[(size_t)b * kParams + p];
    grad[p] += s;
}

// ----------------------------------------------------------------- loss and optimiser
// e = Az - r (t holds A z); partials of (r, r) and (e, e)
__global__ __launch_bounds__(kBlock) void k_gl_err(int N, const float *__restrict__ r, const float *__restrict__ t,
                                                   float *e, double *partials) {
    double rr = 0, ee = 0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock) {
        const float ev = t[i] - r[i];
        e[i] = ev;
        rr += (double)r[i] * (double)r[i];
        ee += (double)ev * (double)ev;
    }
    const double v[2] = {rr, ee};
    store_partials<2>(v, partials);
}
// sc[2] = rr, sc[3] = ee, sc[4] += ee/rr
__global__ __launch_bounds__(kBlock) void k_gl_fin(const double *P, int B, double *sc) {
    const double rr = sum_partials(P, B), ee = sum_partials(P + kPartLd, B);
    if (threadIdx.x == 0) {
        sc[2] = rr;
        sc[3] = ee;
        sc[4] += rr > 0 ? ee / rr : 0.0;
    }
}
// dz = A^T e * (2 / (nb rr))   (t holds A^T e)
__global__ __launch_bounds__(kBlock) void k_gl_dz(int N, const float *__restrict__ t, const double *sc, int nb,
                                                  float *dz) {
    const float f = (float)(2.0 / ((double)nb * sc[2]));
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock)
        dz[i] = t[i] * f;
}
__global__ void k_adam(int n, float *theta, float *m, float *v, const double *grad, double lr, double b1, double b2,
                       double eps, double c1, double c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double g = grad[p];
    const double mm = b1 * m[p] + (1 - b1) * g, vv = b2 * v[p] + (1 - b2) * g * g;
    m[p] = (float)mm;
    v[p] = (float)vv;
    theta[p] = (float)(theta[p] - lr * (mm / c1) / (sqrt(vv / c2) + eps));
}
__global__ void k_zero_d(int n, double *p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline int node_grid(int N) { return (N + kBlock - 1) / kBlock; }
inline int tile_grid(int N) { return (N + 32 * kWavesPerBlock - 1) / (32 * kWavesPerBlock); }

template <typename T>
int dalloc(T **p, size_t n) {
    SENK_HIP(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    return SENK_OK;
}
}  // namespace

// ----------------------------------------------------------------- host helpers
void gnn_graph_free(GnnGraph *g) {
    for (void *p : {(void *)g->s, (void *)g->own_trptr, (void *)g->own_tcind, (void *)g->own_tval, (void *)g->tperm,
                    (void *)g->tptr, (void *)g->ent})
        if (p) (void)hipFree(p);
    *g = GnnGraph{};
}
void gnn_work_free(GnnWork *w) {
    for (void *p : {(void *)w->xin, (void *)w->h1, (void *)w->h2, (void *)w->h3, (void *)w->m2,
                    (void *)w->m3, (void *)w->chat, (void *)w->x, (void *)w->m1, (void *)w->G, (void *)w->G2,
                    (void *)w->P, (void *)w->Q, (void *)w->dchat, (void *)w->z, (void *)w->e, (void *)w->t,
                    (void *)w->tmp, (void *)w->sc, (void *)w->gpart})
        if (p) (void)hipFree(p);
    *w = GnnWork{};
}
int gnn_work_ensure(GnnWork *w, int N, bool train) {
    if (w->N >= N && (w->train || !train)) return SENK_OK;
    gnn_work_free(w);
    w->N = N;
    w->train = train;
    const size_t nf = (size_t)N * gnn::W;
    SENK_TRY(dalloc(&w->xin, N));
    SENK_TRY(dalloc(&w->h1, nf));
    SENK_TRY(dalloc(&w->h2, nf));
    SENK_TRY(dalloc(&w->chat, N));
    SENK_TRY(dalloc(&w->sc, 8));
    if (train) {
        SENK_TRY(dalloc(&w->h3, nf));
        SENK_TRY(dalloc(&w->m2, nf));
        SENK_TRY(dalloc(&w->m3, nf));
        SENK_TRY(dalloc(&w->x, N));
        SENK_TRY(dalloc(&w->m1, N));
        SENK_TRY(dalloc(&w->G, nf));
        SENK_TRY(dalloc(&w->G2, nf));
        SENK_TRY(dalloc(&w->P, nf));
        SENK_TRY(dalloc(&w->Q, nf));
        SENK_TRY(dalloc(&w->dchat, N));
        SENK_TRY(dalloc(&w->z, N));
        SENK_TRY(dalloc(&w->e, N));
        SENK_TRY(dalloc(&w->t, N));
        SENK_TRY(dalloc(&w->tmp, N));
        SENK_TRY(dalloc(&w->gpart, (size_t)gnn::kGradBlocks * gnn::kParams));
    }
    return SENK_OK;
}

// Breadth-first (Cuthill-McKee) order of the tiles of 32 consecutive rows, two tiles
// being neighbours when a row of one has a column in the other; one component after the
// other, each from a pseudo-peripheral tile (two George-Liu steps: restart from a tile
// with fewest entries in the last level while that lengthens the search).  A partial
// last tile (N not a multiple of 32) stays last.
static std::vector<int> tile_bfs_order(int N, const std::vector<int> &rp, const std::vector<int> &ci) {
    const int nt = (N + 31) / 32;
    const int last = (N % 32) ? nt - 1 : -1;
    std::vector<int> order, q, dist((size_t)nt, -1);
    std::vector<char> done((size_t)nt, 0);
    order.reserve((size_t)nt);
    q.reserve((size_t)nt);
    auto entries = [&](int t) { return rp[std::min(N, 32 * t + 32)] - rp[32 * t]; };
    if (last >= 0) done[last] = 1;
    // search from root over tiles not yet ordered; returns the depth, q the visit order
    auto search = [&](int root) {
        q.clear();
        q.push_back(root);
        dist[root] = 0;
        for (size_t h = 0; h < q.size(); h++) {
            const int t = q[h];
            for (int j = rp[32 * t]; j < rp[std::min(N, 32 * t + 32)]; j++) {
                const int c = ci[j] >> 5;
                if (!done[c] && dist[c] < 0) {
                    dist[c] = dist[t] + 1;
                    q.push_back(c);
                }
            }
        }
        return dist[q.back()];
    };
    for (int s0 = 0; s0 < nt; s0++) {
        if (done[s0]) continue;
        int root = s0, depth = -1;
        for (int step = 0; step < 2; step++) {
            const int d = search(root);
            int best = q.back();
            for (size_t h = q.size(); h-- > 0 && dist[q[h]] == d;)
                if (entries(q[h]) < entries(best)) best = q[h];
            for (int v : q) dist[v] = -1;
            if (d <= depth) break;
            depth = d;
            root = best;
        }
        search(root);
        for (int v : q) {
            dist[v] = -1;
            done[v] = 1;
            order.push_back(v);
        }
    }
    if (last >= 0) order.push_back(last);
    return order;
}

// s = |diag|^-1/2, the layers' tile order (breadth-first when `reorder`, else the matrix
// order) and their tiled graph; the transpose (training) on the host, or A's own arrays
// when A is symmetric.
int gnn_graph_build(Ctx *c, int N, const int *drptr, const int *dcind, const float *dval, bool transpose,
                    bool reorder, GnnGraph *g) {
    gnn_graph_free(g);
    g->N = N;
    g->rptr = drptr;
    g->cind = dcind;
    g->val = dval;
    g->trptr = drptr;
    g->tcind = dcind;
    g->tval = dval;
    std::vector<int> rp((size_t)N + 1);
    SENK_HIP(hipMemcpyAsync(rp.data(), drptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    const int nnz = rp[N];
    std::vector<int> ci((size_t)std::max(nnz, 1));
    std::vector<float> va((size_t)std::max(nnz, 1));
    SENK_HIP(hipMemcpyAsync(ci.data(), dcind, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipMemcpyAsync(va.data(), dval, sizeof(float) * nnz, hipMemcpyDeviceToHost, c->stream));
    SENK_HIP(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nnz; j++)
        if (ci[j] < 0 || ci[j] >= N) {
            record_error("learned preconditioner: column index out of range");
            return SENK_ERR_ARG;
        }
    std::vector<float> s((size_t)N);
    for (int i = 0; i < N; i++) {
        double d = 0;
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (ci[j] == i) d = va[j];
        if (d == 0) {
            record_error("learned preconditioner: zero diagonal entry");
            return SENK_ERR_ARG;
        }
        s[i] = (float)(1.0 / std::sqrt(std::fabs(d)));
    }
    // layers' tile T holds matrix rows 32 P[T] + [0, 32); tile of row i: iP[i / 32]
    const int nt = (N + 31) / 32;
    std::vector<int> P, iP;
    if (reorder) {
        P = tile_bfs_order(N, rp, ci);
        iP.assign((size_t)nt, 0);
        for (int t = 0; t < nt; t++) iP[P[t]] = t;
    }
    auto row = [&](int u) { return reorder ? 32 * P[u >> 5] + (u & 31) : u; };
    auto pos = [&](int i) { return reorder ? 32 * iP[i >> 5] + (i & 31) : i; };
    std::vector<int> tb((size_t)nt + 1, 0);
    for (int t = 0; t < nt; t++) {
        int w = 0;
        for (int u = 32 * t; u < std::min(N, 32 * t + 32); u++) w = std::max(w, rp[row(u) + 1] - rp[row(u)]);
        tb[t + 1] = tb[t] + (w + 7) / 8 * 8;
    }
    // padding: the node itself (lanes past N: the tile's first node) with weight 0
    std::vector<GnnEdge> en((size_t)tb[nt] * 32);
    std::vector<float> sn((size_t)N);
    for (int t = 0; t < nt; t++)
        for (int l = 0; l < 32; l++) {
            const int u = 32 * t + l, self = u < N ? u : 32 * t;
            int k = 0;
            if (u < N) {
                const int i = row(u);
                sn[u] = s[i];
                for (int j = rp[i]; j < rp[i + 1]; j++, k++)
                    en[((size_t)tb[t] + k) * 32 + l] = GnnEdge{pos(ci[j]), va[j] * s[ci[j]]};
            }
            for (; k < tb[t + 1] - tb[t]; k++) en[((size_t)tb[t] + k) * 32 + l] = GnnEdge{self, 0.f};
        }
    SENK_TRY(dalloc(&g->s, N));
    SENK_TRY(dalloc(&g->tptr, (size_t)nt + 1));
    SENK_TRY(dalloc(&g->ent, en.size()));
    SENK_HIP(hipMemcpy(g->s, sn.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    SENK_HIP(hipMemcpy(g->tptr, tb.data(), sizeof(int) * (nt + 1), hipMemcpyHostToDevice));
    SENK_HIP(hipMemcpy(g->ent, en.data(), sizeof(GnnEdge) * en.size(), hipMemcpyHostToDevice));
    if (reorder) {
        SENK_TRY(dalloc(&g->tperm, nt));
        SENK_HIP(hipMemcpy(g->tperm, P.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    }
    if (!transpose) return SENK_OK;
    // transpose by counting sort; keep A's arrays when the result is identical
    std::vector<int> tp((size_t)N + 1, 0), tc((size_t)std::max(nnz, 1));
    std::vector<float> tv((size_t)std::max(nnz, 1));
    for (int j = 0; j < nnz; j++) tp[ci[j] + 1]++;
    for (int i = 0; i < N; i++) tp[i + 1] += tp[i];
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < N; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++) {
            const int q = fill[ci[j]]++;
            tc[q] = i;
            tv[q] = va[j];
        }
    bool same = tp == rp;
    for (int j = 0; same && j < nnz; j++) same = tc[j] == ci[j] && tv[j] == va[j];
    if (same) return SENK_OK;
    SENK_TRY(dalloc(&g->own_trptr, (size_t)N + 1));
    SENK_TRY(dalloc(&g->own_tcind, (size_t)nnz));
    SENK_TRY(dalloc(&g->own_tval, (size_t)nnz));
    SENK_HIP(hipMemcpy(g->own_trptr, tp.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    SENK_HIP(hipMemcpy(g->own_tcind, tc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
    SENK_HIP(hipMemcpy(g->own_tval, tv.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
    g->trptr = g->own_trptr;
    g->tcind = g->own_tcind;
    g->tval = g->own_tval;
    return SENK_OK;
}

int gnn_pack(senk_gnn_s *net) {
    k_pack_params<<<(gnn::kParams + 255) / 256, 256, 0, net->c->stream>>>(net->theta, net->theta_r, net->wb);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// z (holding z0) += correction for residual r; keep = store what the backward pass needs
template <typename T>
int gnn_forward(Ctx *c, senk_gnn_s *net, const GnnGraph &g, GnnWork &w, const T *r, const double *in_scale, T *z,
                bool keep) {
    using namespace gnn;
    const int N = g.N;
    const int B = red_grid(N);
    double *P = red_slot(c, kSlot);
    k_gnn_in<T><<<B, kBlock, 0, c->stream>>>(N, r, in_scale, z, g.s, g.tperm, w.xin, P);
    SENK_LAUNCH_CHECK();
    k_gnn_sigma<<<1, kBlock, 0, c->stream>>>(P, B, N, w.sc);
    SENK_LAUNCH_CHECK();
    const float *th = net->theta_r;
    k_gnn_l1<<<node_grid(N), kBlock, 0, c->stream>>>(N, g.tptr, g.ent, g.s, w.xin, w.sc, th, w.h1,
                                                       keep ? w.x : nullptr, keep ? w.m1 : nullptr);
    SENK_LAUNCH_CHECK();
    k_gnn_mc<false><<<tile_grid(N), kBlock, 0, c->stream>>>(N, g.tptr, g.ent, g.s, w.h1, net->wb, th + ob2, w.h2,
                                                             keep ? w.m2 : nullptr, nullptr, nullptr, nullptr);
    SENK_LAUNCH_CHECK();
    k_gnn_mc<true><<<tile_grid(N), kBlock, 0, c->stream>>>(N, g.tptr, g.ent, g.s, w.h2, net->wb + 512, th + ob3,
                                                            keep ? w.h3 : nullptr, keep ? w.m3 : nullptr, th + owo,
                                                            th + obo, w.chat);
    SENK_LAUNCH_CHECK();
    k_gnn_out<T><<<node_grid(N), kBlock, 0, c->stream>>>(N, g.s, g.tperm, w.chat, w.sc, z);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// Backward pass for one residual: dz (device, N) -> gradient added to net->grad.
int gnn_backward(Ctx *c, senk_gnn_s *net, const GnnGraph &g, GnnWork &w, const float *dz) {
    using namespace gnn;
    const int N = g.N, ng = node_grid(N);
    const float *th = net->theta_r;
    k_gb_out<<<ng, kBlock, 0, c->stream>>>(N, dz, g.s, w.sc, th, w.h3, w.dchat, w.G);
    SENK_LAUNCH_CHECK();
    // layer 3: grads, then G2 = dh2 * [h2 > 0]
    k_gb_wmc<<<kGradBlocks, kBlock, 0, c->stream>>>(N, w.h2, w.m3, w.G, oW3s, oW3n, ob3, w.gpart);
    SENK_LAUNCH_CHECK();
    k_gb_pq<<<ng, kBlock, 0, c->stream>>>(N, w.G, th + oW3s, th + oW3n, w.P, w.Q);
    SENK_LAUNCH_CHECK();
    k_gb_agg<<<ng, kBlock, 0, c->stream>>>(N, g.trptr, g.tcind, g.tval, g.s, w.P, w.Q, w.h2, w.G2);
    SENK_LAUNCH_CHECK();
    // layer 2: grads, then G = dh1 * [h1 > 0]
    k_gb_wmc<<<kGradBlocks, kBlock, 0, c->stream>>>(N, w.h1, w.m2, w.G2, oW2s, oW2n, ob2, w.gpart);
    SENK_LAUNCH_CHECK();
    k_gb_pq<<<ng, kBlock, 0, c->stream>>>(N, w.G2, th + oW2s, th + oW2n, w.P, w.Q);
    SENK_LAUNCH_CHECK();
    k_gb_agg<<<ng, kBlock, 0, c->stream>>>(N, g.trptr, g.tcind, g.tval, g.s, w.P, w.Q, w.h1, w.G);
    SENK_LAUNCH_CHECK();
    // layer 1 and output layer
    k_gb_w1o<<<kGradBlocks, kBlock, 0, c->stream>>>(N, w.x, w.m1, w.G, w.h3, w.dchat, w.gpart);
    SENK_LAUNCH_CHECK();
    k_gb_reduce<<<(kParams + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(w.gpart, kGradBlocks, net->grad);
    SENK_LAUNCH_CHECK();
    return SENK_OK;
}

// One residual of a batch of nb: z0 = M^-1 r (FP32 factors) or the given z0 (the result
// of another preconditioner of r), z = z0 + correction, loss term ||A z - r||^2/||r||^2
// added to sc[4], gradient of loss/nb added to grad.
int gnn_step(Ctx *c, senk_gnn_s *net, const GnnGraph &g, GnnWork &w, const MatView &A, const PrecView &M,
             const float *r, const float *z0, int nb) {
    const int N = g.N;
    if (z0) SENK_TRY(launch_copy<float>(c, N, z0, w.z));
    else SENK_TRY(prec_apply<float>(c, M, r, nullptr, w.z, w.tmp, nullptr));
    SENK_TRY(gnn_forward<float>(c, net, g, w, r, nullptr, w.z, true));
    SENK_TRY(launch_spmv_csr<float>(c, N, g.rptr, g.cind, g.val, w.z, w.t, A.avg_row));
    const int B = red_grid(N);
    double *P = red_slot(c, gnn::kSlot);
    k_gl_err<<<B, kBlock, 0, c->stream>>>(N, r, w.t, w.e, P);
    SENK_LAUNCH_CHECK();
    k_gl_fin<<<1, kBlock, 0, c->stream>>>(P, B, w.sc);
    SENK_LAUNCH_CHECK();
    SENK_TRY(launch_spmv_csr<float>(c, N, g.trptr, g.tcind, g.tval, w.e, w.t, A.avg_row));
    k_gl_dz<<<node_grid(N), kBlock, 0, c->stream>>>(N, w.t, w.sc, nb, w.e);  // dz into e
    SENK_LAUNCH_CHECK();
    return gnn_backward(c, net, g, w, w.e);
}

int learned_apply_f32(Ctx *c, const PrecView &M, const float *r, const double *in_scale, float *out,
                      const int *) {
    if (M.learned->off) return SENK_OK;
    senk_gnn_s *net = M.learned->net;
    if (!net || net->g.N != M.N) return SENK_ERR_ARG;
    SENK_TRY(gnn_work_ensure(&net->w, M.N, net->w.train));
    return gnn_forward<float>(c, net, net->g, net->w, r, in_scale, out, false);
}
int learned_apply_f64(Ctx *c, const PrecView &M, const double *r, const double *in_scale, double *out,
                      const int *) {
    if (M.learned->off) return SENK_OK;
    senk_gnn_s *net = M.learned->net;
    if (!net || net->g.N != M.N) return SENK_ERR_ARG;
    SENK_TRY(gnn_work_ensure(&net->w, M.N, net->w.train));
    return gnn_forward<double>(c, net, net->g, net->w, r, in_scale, out, false);
}

// ----------------------------------------------------------------- handle operations
int gnn_create(Ctx *c, senk_gnn_s **out) {
    auto *net = new senk_gnn_s();
    net->c = c;
    net->lp.net = net;
    int st = SENK_OK;
    if ((st = dalloc(&net->theta, gnn::kParams)) || (st = dalloc(&net->theta_r, gnn::kParams)) ||
        (st = dalloc(&net->wb, 1024)) || (st = dalloc(&net->grad, gnn::kParams)) ||
        (st = dalloc(&net->adam_m, gnn::kParams)) || (st = dalloc(&net->adam_v, gnn::kParams))) {
        gnn_destroy(net);
        return st;
    }
    (void)hipMemsetAsync(net->theta, 0, sizeof(float) * gnn::kParams, c->stream);
    (void)hipMemsetAsync(net->adam_m, 0, sizeof(float) * gnn::kParams, c->stream);
    (void)hipMemsetAsync(net->adam_v, 0, sizeof(float) * gnn::kParams, c->stream);
    if ((st = gnn_pack(net))) {
        gnn_destroy(net);
        return st;
    }
    *out = net;
    return SENK_OK;
}
void gnn_destroy(senk_gnn_s *net) {
    if (!net) return;
    if (net->c) (void)hipStreamSynchronize(net->c->stream);
    for (void *p : {(void *)net->theta, (void *)net->theta_r, (void *)net->wb, (void *)net->grad,
                    (void *)net->adam_m, (void *)net->adam_v})
        if (p) (void)hipFree(p);
    gnn_graph_free(&net->g);
    gnn_work_free(&net->w);
    delete net;
}
int gnn_set_params(senk_gnn_s *net, const float *h) {
    SENK_HIP(hipMemcpyAsync(net->theta, h, sizeof(float) * gnn::kParams, hipMemcpyHostToDevice, net->c->stream));
    SENK_TRY(gnn_pack(net));
    SENK_HIP(hipStreamSynchronize(net->c->stream));
    return SENK_OK;
}
int gnn_get_params(senk_gnn_s *net, float *h) {
    SENK_HIP(hipMemcpyAsync(h, net->theta, sizeof(float) * gnn::kParams, hipMemcpyDeviceToHost, net->c->stream));
    SENK_HIP(hipStreamSynchronize(net->c->stream));
    return SENK_OK;
}
int gnn_bind(senk_gnn_s *net, int N, const int *drptr, const int *dcind, const float *dval) {
    if (!dval) return SENK_ERR_ARG;  // the network aggregates over the FP32 values
    SENK_TRY(gnn_graph_build(net->c, N, drptr, dcind, dval, false, true, &net->g));
    return gnn_work_ensure(&net->w, N, net->w.train);
}
LearnedPrec *gnn_learned(senk_gnn_s *net) { return &net->lp; }
int gnn_apply_f32(senk_gnn_s *net, const float *r, float *z) {
    if (net->g.N <= 0) return SENK_ERR_ARG;
    return gnn_forward<float>(net->c, net, net->g, net->w, r, nullptr, z, false);
}

// Mean loss over nvec residuals (device, FP32, rows of N) and its gradient (host, FP64).
int gnn_loss_grad(senk_gnn_s *net, const MatView &A, const PrecView &M, const float *R, const float *Z, int nvec,
                  double *loss, double *grad_host) {
    Ctx *c = net->c;
    if (!A.v32 || nvec <= 0) return SENK_ERR_ARG;
    GnnGraph g;
    int st = gnn_graph_build(c, A.N, A.rptr, A.cind, A.v32, true, false, &g);
    if (st == SENK_OK) st = gnn_work_ensure(&net->w, A.N, true);
    if (st == SENK_OK) {
        k_zero_d<<<(gnn::kParams + 255) / 256, 256, 0, c->stream>>>(gnn::kParams, net->grad);
        k_zero_d<<<1, 64, 0, c->stream>>>(8, net->w.sc);
        for (int b = 0; b < nvec && st == SENK_OK; b++)
            st = gnn_step(c, net, g, net->w, A, M, R + (size_t)b * A.N, Z ? Z + (size_t)b * A.N : nullptr, nvec);
    }
    double sc[8] = {0};
    if (st == SENK_OK && (hipMemcpyAsync(sc, net->w.sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                          (grad_host && hipMemcpyAsync(grad_host, net->grad, sizeof(double) * gnn::kParams,
                                                       hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
                          hipStreamSynchronize(c->stream) != hipSuccess))
        st = SENK_ERR_HIP;
    if (loss) *loss = sc[4] / nvec;
    gnn_graph_free(&g);
    return st;
}

// Adam over the residual sets of several matrices: every step takes `batch` residuals of
// one matrix; matrices and residual order are shuffled per epoch from `seed`.
int gnn_train(senk_gnn_s *net, int nmat, const MatView *As, const PrecView *Ms, const float *const *Rs,
              const float *const *Zs, const int *nvecs, int epochs, int batch, double lr, unsigned seed,
              double *loss_hist, int verbose) {
    Ctx *c = net->c;
    if (nmat <= 0 || epochs <= 0 || batch <= 0) return SENK_ERR_ARG;
    std::vector<GnnGraph> gs((size_t)nmat);
    int st = SENK_OK, maxN = 0;
    for (int k = 0; k < nmat && st == SENK_OK; k++) {
        if (!As[k].v32 || (!Ms[k].l32 && !(Zs && Zs[k]))) st = SENK_ERR_ARG;
        else st = gnn_graph_build(c, As[k].N, As[k].rptr, As[k].cind, As[k].v32, true, false, &gs[k]);
        maxN = std::max(maxN, As[k].N);
    }
    if (st == SENK_OK) st = gnn_work_ensure(&net->w, maxN, true);
    // (matrix, first residual) pairs of every step of an epoch
    std::vector<std::pair<int, int>> steps;
    for (int k = 0; k < nmat; k++)
        for (int b = 0; b < nvecs[k]; b += batch) steps.emplace_back(k, b);
    std::mt19937 rng(seed);
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    for (int ep = 0; ep < epochs && st == SENK_OK; ep++) {
        std::shuffle(steps.begin(), steps.end(), rng);
        k_zero_d<<<1, 64, 0, c->stream>>>(8, net->w.sc);
        long long nterms = 0;
        for (const auto &sp : steps) {
            const int k = sp.first, nb = std::min(batch, nvecs[k] - sp.second);
            k_zero_d<<<(gnn::kParams + 255) / 256, 256, 0, c->stream>>>(gnn::kParams, net->grad);
            for (int b = 0; b < nb && st == SENK_OK; b++) {
                const size_t off = (size_t)(sp.second + b) * As[k].N;
                st = gnn_step(c, net, gs[k], net->w, As[k], Ms[k], Rs[k] + off,
                              Zs && Zs[k] ? Zs[k] + off : nullptr, nb);
            }
            nterms += nb;
            if (st) break;
            net->adam_t++;
            const double c1 = 1 - std::pow(b1, (double)net->adam_t), c2 = 1 - std::pow(b2, (double)net->adam_t);
            k_adam<<<(gnn::kParams + 255) / 256, 256, 0, c->stream>>>(gnn::kParams, net->theta, net->adam_m,
                                                                      net->adam_v, net->grad, lr, b1, b2, eps, c1, c2);
            if (hipGetLastError() != hipSuccess || (st = gnn_pack(net))) {
                if (!st) st = SENK_ERR_HIP;
                break;
            }
        }
        double sc[8] = {0};
        if (st == SENK_OK && (hipMemcpyAsync(sc, net->w.sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream) !=
                                  hipSuccess ||
                              hipStreamSynchronize(c->stream) != hipSuccess))
            st = SENK_ERR_HIP;
        if (st == SENK_OK && loss_hist) loss_hist[ep] = sc[4] / std::max(1LL, nterms);
        if (st == SENK_OK && verbose) {
            std::printf("epoch %d loss %.6e\n", ep, sc[4] / std::max(1LL, nterms));
            std::fflush(stdout);
        }
    }
    for (auto &g : gs) gnn_graph_free(&g);
    return st;
}

}  // namespace senk_amd
