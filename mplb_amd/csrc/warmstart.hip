// warmstart.hip -- kernels of the learned warm start (a small U-Net that
// predicts the steady cavity flow from the obstacle mask and Re): 3x3
// convolutions as implicit GEMMs on the gfx950 matrix cores
// (v_mfma_f32_32x32x8_bf16: BF16 operands, FP32 accumulation), their data
// and weight gradients, and the element-wise layers, loss and optimiser.
// C ABI in include/mplbws.h.  No reference counterpart: the reference
// starts every run from its initial macros.
//
// Tensors are FP32 in HBM, channels last: x[n][y][x][c].  Operands are
// rounded to BF16 (nearest even) as they are loaded into the MFMA
// fragments.  Convolution weights are laid out [tap = 3*ky + kx][cin][cout].
#include "../../include/mplbws.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;  // waves per workgroup

// FP32 -> BF16, round to nearest even (finite inputs), in integer
// arithmetic: measured faster here than the packed hardware conversion
// (v_cvt_pk_bf16_f32) in these staging loops
__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
// two at once, a in the low half
__device__ __forceinline__ uint32_t to_bf16x2(float a, float b) {
    return (uint32_t)(uint16_t)to_bf16(a) | ((uint32_t)(uint16_t)to_bf16(b) << 16);
}
__device__ __forceinline__ uint2 to_bf16x4(const float4& f) {
    return uint2{to_bf16x2(f.x, f.y), to_bf16x2(f.z, f.w)};
}

// Forward 3x3 convolution as an implicit GEMM, out[m][co] = sum_{tap, ci}
// in[shift_tap(m)][ci] * w[tap][ci][co], zero padding at the image edges.
//
// A workgroup (4 waves) owns a tile of TP = 128 WM output pixels -- R whole
// rows of W0 = min(wd, 128) columns -- and CG = 32 WN output channels.
// Input channels go in chunks of kCC: the chunk's input tile with its
// one-pixel halo is converted to BF16 once into LDS (zeros outside the
// image), and so are the chunk's weights, in the MFMA operand layout; the
// waves then run the 9 taps from LDS only.  Wave v takes pixels
// [32 WM v, 32 WM (v + 1)) of the tile and all CG channels: WM x WN
// accumulators of 32x32, each A fragment feeding WN of them and each B
// fragment WM.  WM = 2 for large images (fewer, fuller workgroups), 1 where
// that leaves too few workgroups; WN = 1 for layers of at most 32 output
// channels.
//
// MFMA 32x32x8 (BF16 in, FP32 accumulate) lane layout: lane l holds
// A[l%32][4(l/32) .. +3] and B[4(l/32) .. +3][l%32]; the accumulator
// element 4i+j of lane l is C[8i + 4(l/32) + j][l%32].
constexpr int kCC = 32;              // input channels per LDS chunk
constexpr int kPixStride = kCC + 4;  // BF16 elements per staged pixel (the
                                     // 8-byte pad spreads a wave's reads
                                     // over all LDS banks)
constexpr int kMaxW0 = 128;

struct ConvTile {
    int W0, R, tiles_x, tiles_y;
};

__host__ __device__ inline ConvTile conv_tile(int h, int wd, int tp) {
    ConvTile t;
    t.W0 = wd < kMaxW0 ? wd : kMaxW0;
    t.R = tp / t.W0;
    if (t.R < 1) t.R = 1;
    if (t.R > h) t.R = h;
    t.tiles_x = (wd + t.W0 - 1) / t.W0;
    t.tiles_y = (h + t.R - 1) / t.R;
    return t;
}

// pixels staged for a tile of tp pixels: the largest (R + 2)(W0 + 2) over
// the widths; the host refuses shapes above it
__host__ __device__ constexpr int conv_max_staged(int tp) { return tp == 256 ? 520 : 390; }

// q = e / d, r = e - q d for 0 <= e < 2^22 and 1 <= d <= 2^10 through a
// float reciprocal (error below one unit before the correction)
struct FastDiv {
    float inv;
    int d;
};
__device__ __forceinline__ FastDiv fast_div(int d) { return FastDiv{1.0f / (float)d, d}; }
__device__ __forceinline__ int fdiv(int e, const FastDiv& f, int* r) {
    int q = (int)((float)e * f.inv);
    int rr = e - q * f.d;
    if (rr < 0) { q--; rr += f.d; }
    if (rr >= f.d) { q++; rr -= f.d; }
    *r = rr;
    return q;
}

template <int WM, int WN>
__global__ __launch_bounds__(64 * kWaves) void k_conv_fwd(const float* __restrict__ in,
                                                          const float* __restrict__ w,
                                                          const float* __restrict__ bias,
                                                          float* __restrict__ out, int n,
                                                          int h, int wd, int cin, int cout,
                                                          int relu) {
    constexpr int TP = 128 * WM, CG = 32 * WN;
    __shared__ __attribute__((aligned(16))) short s_in[conv_max_staged(TP) * kPixStride];
    // B fragments: [tap][kstep][g][co 0 .. CG-1][4]
    __shared__ __attribute__((aligned(16))) short s_w[9 * (kCC / 8) * 2 * CG * 4];
    const ConvTile T = conv_tile(h, wd, TP);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row = lane & 31, g = lane >> 5;
    const int tiles = T.tiles_x * T.tiles_y;
    const int pn = (int)blockIdx.x / tiles;
    const int t = (int)blockIdx.x % tiles;
    const int y0 = (t / T.tiles_x) * T.R, x0 = (t % T.tiles_x) * T.W0;
    const int co0 = (int)blockIdx.y * CG;
    const int SW = T.W0 + 2, SH = T.R + 2;
    const FastDiv dW0 = fast_div(T.W0);
    // this lane's A rows (pixels) in its WM pixel tiles; a lane past the
    // tile reads pixel (0, 0) instead (its results are never stored)
    int pbase[WM];
#pragma unroll
    for (int mt = 0; mt < WM; mt++) {
        int tx;
        int ty = fdiv(wave * 32 * WM + mt * 32 + row, dW0, &tx);
        if (ty >= T.R) ty = tx = 0;
        pbase[mt] = ((ty + 1) * SW + tx + 1) * kPixStride + 4 * g;
    }
    f32x16 acc[WM][WN] = {};
    const float* img = in + (int64_t)pn * h * wd * cin;
    for (int c0 = 0; c0 < cin; c0 += kCC) {
        const int cc = cin - c0 < kCC ? cin - c0 : kCC;  // a multiple of 8
        const int q8 = cc / 8;
        __syncthreads();  // the previous chunk's reads are done
        // input tile + halo in BF16, channels c0 .. c0 + cc: item e = one
        // staged pixel's 8 channels (consecutive items: consecutive channel
        // groups, then pixels); each thread issues the loads of kBatch
        // items before converting any, so they are in flight together
        {
            const int total = SH * SW * q8;
            const FastDiv dq8 = fast_div(q8), dsw = fast_div(SW);
            constexpr int kBatch = 4;
            for (int e0 = threadIdx.x; e0 < total; e0 += 64 * kWaves * kBatch) {
                float4 f[kBatch][2];
                int slot[kBatch];
#pragma unroll
                for (int k = 0; k < kBatch; k++) {
                    const int e = e0 + k * 64 * kWaves;
                    f[k][0] = f[k][1] = float4{0.f, 0.f, 0.f, 0.f};
                    slot[k] = -1;
                    if (e < total) {
                        int c8, sx;
                        const int sp = fdiv(e, dq8, &c8);
                        const int sy = fdiv(sp, dsw, &sx);
                        const int yy = y0 + sy - 1, xx = x0 + sx - 1;
                        slot[k] = sp * kPixStride + 8 * c8;
                        if (yy >= 0 && yy < h && xx >= 0 && xx < wd) {
                            const float* src = img + ((int64_t)yy * wd + xx) * cin + c0 + 8 * c8;
                            f[k][0] = *reinterpret_cast<const float4*>(src);
                            f[k][1] = *reinterpret_cast<const float4*>(src + 4);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < kBatch; k++) {
                    if (slot[k] < 0) continue;
                    *reinterpret_cast<uint2*>(&s_in[slot[k]]) = to_bf16x4(f[k][0]);
                    *reinterpret_cast<uint2*>(&s_in[slot[k] + 4]) = to_bf16x4(f[k][1]);
                }
            }
        }
        // weights of the chunk in B-fragment order, zero past cout: thread
        // -> four output channels co4 .. co4 + 3 (one 16-byte load when
        // they lie inside cout and cout is a multiple of 4) and the (tap,
        // ci) pairs j = tid / (CG / 4) + k (256 / (CG / 4)); every load of
        // the chunk is issued before any is converted, so they are in
        // flight together
        {
            constexpr int groups = CG / 4;
            constexpr int step = 64 * kWaves / groups;
            constexpr int kMax = (9 * kCC + step - 1) / step;  // loads per thread
            const int co4 = ((int)threadIdx.x % groups) * 4;
            const int oc = co0 + co4;
            const int pairs = 9 * cc;
            const int j0 = (int)threadIdx.x / groups;
            const FastDiv dcc = fast_div(cc);
            const bool vec = (cout & 3) == 0 && oc + 3 < cout;
            float4 v[kMax];
#pragma unroll
            for (int k = 0; k < kMax; k++) {
                const int j = j0 + k * step;
                v[k] = float4{0.f, 0.f, 0.f, 0.f};
                if (j < pairs && oc < cout) {
                    int ci;
                    const int tap = fdiv(j, dcc, &ci);
                    const float* src = w + ((int64_t)tap * cin + c0 + ci) * cout + oc;
                    if (vec) {
                        v[k] = *reinterpret_cast<const float4*>(src);
                    } else {
                        v[k].x = src[0];
                        if (oc + 1 < cout) v[k].y = src[1];
                        if (oc + 2 < cout) v[k].z = src[2];
                        if (oc + 3 < cout) v[k].w = src[3];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kMax; k++) {
                const int j = j0 + k * step;
                if (j >= pairs) break;
                int ci;
                const int tap = fdiv(j, dcc, &ci);
                const int ks = ci >> 3, gg = (ci >> 2) & 1, tt = ci & 3;
                short* dst = &s_w[((((tap * (kCC / 8) + ks) * 2 + gg) * CG) + co4) * 4 + tt];
                dst[0] = to_bf16(v[k].x);
                dst[4] = to_bf16(v[k].y);
                dst[8] = to_bf16(v[k].z);
                dst[12] = to_bf16(v[k].w);
            }
        }
        __syncthreads();
        for (int tap = 0; tap < 9; tap++) {
            const int shift = ((tap / 3 - 1) * SW + (tap % 3 - 1)) * kPixStride;
            for (int ks = 0; ks < q8; ks++) {
                bf16x4 a[WM], b[WN];
#pragma unroll
                for (int mt = 0; mt < WM; mt++)
                    a[mt] = *reinterpret_cast<const bf16x4*>(&s_in[pbase[mt] + shift + ks * 8]);
#pragma unroll
                for (int nt = 0; nt < WN; nt++)
                    b[nt] = *reinterpret_cast<const bf16x4*>(
                        &s_w[((((tap * (kCC / 8) + ks) * 2 + g) * CG) + nt * 32 + row) * 4]);
#pragma unroll
                for (int mt = 0; mt < WM; mt++)
#pragma unroll
                    for (int nt = 0; nt < WN; nt++)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[mt], b[nt],
                                                                                acc[mt][nt], 0, 0, 0);
            }
        }
    }
    // epilogue: bias, relu, FP32 store (32 consecutive channels per row)
#pragma unroll
    for (int nt = 0; nt < WN; nt++) {
        const int co = co0 + nt * 32 + row;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int mt = 0; mt < WM; mt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int tx1;
                const int p1 = wave * 32 * WM + mt * 32 + 8 * i + 4 * g;
                const int ty1 = fdiv(p1, dW0, &tx1);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int tx = tx1 + j, ty = ty1;
                    if (tx >= T.W0) ty = fdiv(p1 + j, dW0, &tx);
                    if (ty >= T.R || y0 + ty >= h || x0 + tx >= wd) continue;
                    float v = acc[mt][nt][4 * i + j] + bv;
                    if (relu) v = v > 0.f ? v : 0.f;
                    out[(((int64_t)pn * h + y0 + ty) * wd + x0 + tx) * cout + co] = v;
                }
            }
    }
}

// weights of the data gradient: wt[tap][co][ci] = w[8 - tap][ci][co]
__global__ void k_flip_weights(const float* __restrict__ w, float* __restrict__ wt, int cin,
                               int cout) {
    const int64_t total = 9LL * cin * cout;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const int co = (int)((i / cin) % cout);
        const int tap = (int)(i / ((int64_t)cin * cout));
        wt[i] = w[((int64_t)(8 - tap) * cin + ci) * cout + co];
    }
}

// Pixel chunks of the weight gradient: whole rows of one image, about
// MPLBWS_WGRAD_CHUNK pixels each, so that a chunk's pixel coordinates need
// no division.
__host__ __device__ inline int wgrad_rows(int h, int wd) {
    int r = MPLBWS_WGRAD_CHUNK / wd;
    if (r < 1) r = 1;
    if (r > h) r = h;
    return r;
}

// Weight gradient partials of one chunk: a wave computes a 32 (rows k =
// tap*cin + ci) x 32 (co) tile of sum_m A[m][k] dout[m][co] over the
// chunk's pixels, 8 per product (MFMA reduction dimension = pixels); the
// waves take the chunk's rows in turn and their tiles are summed through
// LDS in wave order.  partial[chunk][k][co].  The workgroups of the first
// k tile also sum the dout values they load (FP32, fixed order) into the
// bias-gradient partial bpart[chunk][co] when bpart is given.
__global__ __launch_bounds__(64 * kWaves) void k_conv_wgrad(const float* __restrict__ in,
                                                            const float* __restrict__ dout,
                                                            float* __restrict__ partial,
                                                            float* __restrict__ bpart,
                                                            int h, int wd, int cin, int cout,
                                                            int rows) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int K = 9 * cin;
    const int k0 = blockIdx.x * 32, co0 = blockIdx.y * 32;
    const int per_image = (h + rows - 1) / rows;
    const int pn = (int)blockIdx.z / per_image;
    const int r0 = ((int)blockIdx.z % per_image) * rows;
    const int r1 = r0 + rows < h ? r0 + rows : h;
    const int row = lane & 31, g = lane >> 5;
    const int k = k0 + row;  // A row of this lane
    const bool kvalid = k < K;
    const int tap = kvalid ? k / cin : 0, ci = kvalid ? k % cin : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int co = co0 + row;  // B column of this lane
    const bool cvalid = co < cout;
    const float* inb = in + (int64_t)pn * h * wd * cin;
    const float* db = dout + (int64_t)pn * h * wd * cout;
    f32x16 acc = {};
    float bsum = 0.f;
    for (int py = r0 + wave; py < r1; py += kWaves) {
        const int yy = py + dy;
        const bool yin = kvalid && yy >= 0 && yy < h;
        for (int x0 = 0; x0 < wd; x0 += 8) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int px = x0 + 4 * g + t;
                const int xx = px + dx;
                float av = 0.f, bv = 0.f;
                if (px < wd) {
                    if (yin && xx >= 0 && xx < wd) av = inb[((int64_t)yy * wd + xx) * cin + ci];
                    if (cvalid) bv = db[((int64_t)py * wd + px) * cout + co];
                }
                bsum += bv;
                a[t] = to_bf16(av);
                b[t] = to_bf16(bv);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    __shared__ float bred[kWaves][2][32];
    if (bpart && blockIdx.x == 0) {
        bred[wave][g][row] = bsum;
        __syncthreads();
        if (threadIdx.x < 32 && co0 + (int)threadIdx.x < cout) {
            float s = 0.f;
            for (int v = 0; v < kWaves; v++) s += bred[v][0][threadIdx.x] + bred[v][1][threadIdx.x];
            bpart[(int64_t)blockIdx.z * cout + co0 + threadIdx.x] = s;
        }
    }
    // sum the waves' tiles through LDS, then write this chunk's partial
    __shared__ float red[kWaves][32][33];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) red[wave][8 * i + 4 * g + j][row] = acc[4 * i + j];
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += 64 * kWaves) {
        const int r = e / 32, cc = e % 32;
        float s = 0.f;
        for (int v = 0; v < kWaves; v++) s += red[v][r][cc];
        const int kk = k0 + r, oc = co0 + cc;
        if (kk < K && oc < cout) partial[((int64_t)blockIdx.z * K + kk) * cout + oc] = s;
    }
}

// dw = sum over chunks (chunk order: deterministic)
// dst[i] = sum over the chunks of partial[c][i]: a workgroup takes 64
// consecutive i, its four waves the chunks c = w, w + 4, ... (so each wave
// has a quarter of the dependent adds), and the waves' sums are added in
// wave order through LDS: a fixed order, reproducible sums.
constexpr int kSumWaves = 4;
__global__ __launch_bounds__(64 * kSumWaves) void k_sum_partials(const float* __restrict__ partial,
                                                                 float* __restrict__ dst,
                                                                 int64_t count, int nchunks) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    float s = 0.f;
    if (i < count)
        for (int c = w; c < nchunks; c += kSumWaves) s += partial[(int64_t)c * count + i];
    __shared__ float red[kSumWaves][64];
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && i < count) {
        float t = red[0][lane];
        for (int v = 1; v < kSumWaves; v++) t += red[v][lane];
        dst[i] = t;
    }
}

__global__ void k_relu_bwd(const float* __restrict__ dout, const float* __restrict__ out,
                           float* __restrict__ din, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        din[i] = out[i] > 0.f ? dout[i] : 0.f;
}

// 2x2 average pooling and its adjoint (h, wd: the fine size)
__global__ void k_pool_fwd(const float* __restrict__ in, float* __restrict__ out, int n, int h,
                           int wd, int c) {
    const int ho = h / 2, wo = wd / 2;
    const int64_t total = (int64_t)n * ho * wo * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        const int x = (int)(p % wo), y = (int)((p / wo) % ho);
        const int64_t b = p / ((int64_t)wo * ho);
        const float* s = in + ((b * h + 2 * y) * wd + 2 * x) * c + ch;
        out[i] = 0.25f * ((s[0] + s[c]) + (s[(int64_t)wd * c] + s[(int64_t)wd * c + c]));
    }
}

__global__ void k_pool_bwd(const float* __restrict__ dout, float* __restrict__ din, int n,
                           int h, int wd, int c) {
    const int64_t total = (int64_t)n * h * wd * c;
    const int ho = h / 2, wo = wd / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        const int x = (int)(p % wd), y = (int)((p / wd) % h);
        const int64_t b = p / ((int64_t)wd * h);
        din[i] = 0.25f * dout[((b * ho + y / 2) * wo + x / 2) * c + ch];
    }
}

// nearest 2x upsampling (h, wd: the coarse size) into channels [c0, c0+c)
// of a tensor with ctot channels, and its adjoint
__global__ void k_up_fwd(const float* __restrict__ in, float* __restrict__ out, int n, int h,
                         int wd, int c, int ctot, int c0) {
    const int H = 2 * h, W = 2 * wd;
    const int64_t total = (int64_t)n * H * W * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const int64_t b = p / ((int64_t)W * H);
        out[p * ctot + c0 + ch] = in[((b * h + y / 2) * wd + x / 2) * c + ch];
    }
}

__global__ void k_up_bwd(const float* __restrict__ dout, float* __restrict__ din, int n, int h,
                         int wd, int c, int ctot, int c0) {
    const int W = 2 * wd;
    const int64_t total = (int64_t)n * h * wd * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        const int x = (int)(p % wd), y = (int)((p / wd) % h);
        const int64_t b = p / ((int64_t)wd * h);
        const int64_t f = ((b * 2 * h + 2 * y) * W + 2 * x) * ctot + c0 + ch;
        const int64_t r = (int64_t)W * ctot;
        din[i] = (dout[f] + dout[f + ctot]) + (dout[f + r] + dout[f + r + ctot]);
    }
}

// copy channels [0, c) of src (c channels) into channels [c0, c0 + c) of
// dst (ctot channels) (dir = 0), or back (dir = 1), or add them back (dir = 2)
__global__ void k_channels(const float* __restrict__ a, float* __restrict__ b, int64_t pixels,
                           int c, int ctot, int c0, int dir) {
    const int64_t total = pixels * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        if (dir == 0) b[p * ctot + c0 + ch] = a[i];
        else if (dir == 1) b[i] = a[p * ctot + c0 + ch];
        else b[i] += a[p * ctot + c0 + ch];
    }
}

// dst[p][ch] = src[p][ch] on pixels with mask[p] == 0 and ch < cu, else 0
__global__ void k_mask_channels(const float* src, const float* __restrict__ mask, float* dst,
                                int64_t pixels, int c, int cu) {
    const int64_t total = pixels * c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int64_t p = i / c;
        dst[i] = (ch < cu && mask[p] == 0.f) ? src[i] : 0.f;
    }
}

// dst[i][:] = src[idx[i]][:] for rows of `elems` floats (a mini-batch
// from a device-resident training set)
__global__ void k_gather(const float* __restrict__ src, float* __restrict__ dst,
                         const int32_t* __restrict__ idx, int count, int64_t elems) {
    const int64_t total = (int64_t)count * elems;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / elems);
        dst[i] = src[(int64_t)idx[r] * elems + (i - (int64_t)r * elems)];
    }
}

// masked mean squared error over channels [0, cu) of pred (ctot channels):
// per-block partial sums of the loss; grad = 2 (pred - target) / count on
// fluid pixels (mask == 0), zero elsewhere and in channels >= cu
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      const float* __restrict__ mask, float* __restrict__ grad,
                      float* __restrict__ partial, int64_t pixels, int cu, int ctot,
                      float inv_count) {
    __shared__ float red[256];
    float s = 0.f;
    const int64_t total = pixels * ctot;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % ctot);
        const int64_t p = i / ctot;
        float gv = 0.f;
        if (ch < cu && mask[p] == 0.f) {
            const float d = pred[i] - target[p * cu + ch];
            s += d * d;
            gv = 2.f * d * inv_count;
        }
        grad[i] = gv;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

__global__ void k_zero(float* __restrict__ p, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

int grid_for(int64_t total) {
    int64_t g = (total + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}

int status() { return hipGetLastError() == hipSuccess ? 0 : MPLBWS_ERR_HIP; }

}  // namespace

extern "C" {

int mplbws_conv3x3_fwd(const float* in, const float* w, const float* bias, float* out, int n,
                       int h, int wd, int cin, int cout, int relu) {
    if (!in || !w || !out || n <= 0 || h <= 0 || wd <= 0 || cin <= 0 || cout <= 0 ||
        cin % 8 != 0)
        return MPLBWS_ERR_ARG;
    // WM = 2 (256-pixel tiles) unless that gives fewer than 1024
    // workgroups; WN = 1 for at most 32 output channels
    const int wn = cout > 32 ? 2 : 1;
    const int64_t groups = (cout + 32 * wn - 1) / (32 * wn);
    const ConvTile T2 = conv_tile(h, wd, 256);
    const int wm = (int64_t)n * T2.tiles_x * T2.tiles_y * groups >= 1024 &&
                           (T2.R + 2) * (T2.W0 + 2) <= conv_max_staged(256)
                       ? 2
                       : 1;
    const ConvTile T = conv_tile(h, wd, 128 * wm);
    const int64_t wgs = (int64_t)n * T.tiles_x * T.tiles_y;
    if (wgs >= ((int64_t)1 << 31) || (T.R + 2) * (T.W0 + 2) > conv_max_staged(128 * wm))
        return MPLBWS_ERR_ARG;
    const dim3 grd((unsigned)wgs, (unsigned)groups), blk(64 * kWaves);
    if (wm == 2 && wn == 2)
        hipLaunchKernelGGL((k_conv_fwd<2, 2>), grd, blk, 0, 0, in, w, bias, out, n, h, wd, cin,
                           cout, relu);
    else if (wm == 2)
        hipLaunchKernelGGL((k_conv_fwd<2, 1>), grd, blk, 0, 0, in, w, bias, out, n, h, wd, cin,
                           cout, relu);
    else if (wn == 2)
        hipLaunchKernelGGL((k_conv_fwd<1, 2>), grd, blk, 0, 0, in, w, bias, out, n, h, wd, cin,
                           cout, relu);
    else
        hipLaunchKernelGGL((k_conv_fwd<1, 1>), grd, blk, 0, 0, in, w, bias, out, n, h, wd, cin,
                           cout, relu);
    return status();
}

int mplbws_flip_weights(const float* w, float* wt, int cin, int cout) {
    if (!w || !wt || cin <= 0 || cout <= 0) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_flip_weights, dim3(grid_for(9LL * cin * cout)), dim3(256), 0, 0, w, wt,
                       cin, cout);
    return status();
}

int64_t mplbws_wgrad_workspace(int n, int h, int wd, int cin, int cout) {
    if (n <= 0 || h <= 0 || wd <= 0) return 0;
    const int rows = wgrad_rows(h, wd);
    const int64_t chunks = (int64_t)n * ((h + rows - 1) / rows);
    return chunks * 9LL * cin * cout + chunks * cout;
}

int mplbws_conv3x3_wgrad(const float* in, const float* dout, float* dw, float* dbias,
                         float* workspace, int n, int h, int wd, int cin, int cout) {
    if (!in || !dout || !dw || !workspace || n <= 0 || h <= 0 || wd <= 0 || cin <= 0 ||
        cout <= 0)
        return MPLBWS_ERR_ARG;
    const int rows = wgrad_rows(h, wd);
    const int64_t chunks = (int64_t)n * ((h + rows - 1) / rows);
    if (chunks > 65535) return MPLBWS_ERR_ARG;
    const int K = 9 * cin;
    const dim3 grd((unsigned)((K + 31) / 32), (unsigned)((cout + 31) / 32), (unsigned)chunks);
    const int64_t cnt = (int64_t)K * cout;
    float* bp = dbias ? workspace + chunks * cnt : nullptr;
    hipLaunchKernelGGL(k_conv_wgrad, grd, dim3(64 * kWaves), 0, 0, in, dout, workspace, bp, h,
                       wd, cin, cout, rows);
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((cnt + 63) / 64)), dim3(64 * kSumWaves), 0,
                       0, workspace, dw, cnt, (int)chunks);
    if (dbias)
        hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((cout + 63) / 64)),
                           dim3(64 * kSumWaves), 0, 0, bp, dbias, (int64_t)cout, (int)chunks);
    return status();
}

int mplbws_relu_bwd(const float* dout, const float* out, float* din, int64_t count) {
    if (!dout || !out || !din || count < 0) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(count)), dim3(256), 0, 0, dout, out, din, count);
    return status();
}

int mplbws_pool2_fwd(const float* in, float* out, int n, int h, int wd, int c) {
    if (!in || !out || h % 2 || wd % 2) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_pool_fwd, dim3(grid_for((int64_t)n * h * wd * c / 4)), dim3(256), 0, 0,
                       in, out, n, h, wd, c);
    return status();
}

int mplbws_pool2_bwd(const float* dout, float* din, int n, int h, int wd, int c) {
    if (!dout || !din || h % 2 || wd % 2) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for((int64_t)n * h * wd * c)), dim3(256), 0, 0, dout,
                       din, n, h, wd, c);
    return status();
}

int mplbws_up2_fwd(const float* in, float* out, int n, int h, int wd, int c, int ctot, int c0) {
    if (!in || !out || c0 < 0 || c0 + c > ctot) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_up_fwd, dim3(grid_for((int64_t)n * 4 * h * wd * c)), dim3(256), 0, 0, in,
                       out, n, h, wd, c, ctot, c0);
    return status();
}

int mplbws_up2_bwd(const float* dout, float* din, int n, int h, int wd, int c, int ctot, int c0) {
    if (!dout || !din || c0 < 0 || c0 + c > ctot) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_up_bwd, dim3(grid_for((int64_t)n * h * wd * c)), dim3(256), 0, 0, dout,
                       din, n, h, wd, c, ctot, c0);
    return status();
}

int mplbws_channels(const float* src, float* dst, int64_t pixels, int c, int ctot, int c0,
                    int direction) {
    if (!src || !dst || c0 < 0 || c0 + c > ctot) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_channels, dim3(grid_for(pixels * c)), dim3(256), 0, 0, src, dst, pixels,
                       c, ctot, c0, direction);
    return status();
}

int mplbws_mask_channels(const float* src, const float* mask, float* dst, int64_t pixels, int c,
                         int cu) {
    if (!src || !mask || !dst || pixels < 0 || c <= 0 || cu < 0) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_mask_channels, dim3(grid_for(pixels * c)), dim3(256), 0, 0, src, mask,
                       dst, pixels, c, cu);
    return status();
}

int mplbws_gather(const float* src, float* dst, const int32_t* idx, int count,
                  int64_t elems) {
    if (!src || !dst || !idx || count < 0 || elems < 0) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)count * elems)), dim3(256), 0, 0, src,
                       dst, idx, count, elems);
    return status();
}

int mplbws_mse(const float* pred, const float* target, const float* mask, float* grad,
               float* partial, int npartial, int64_t pixels, int cu, int ctot, float inv_count) {
    if (!pred || !target || !mask || !grad || !partial || npartial <= 0 || cu > ctot)
        return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_mse, dim3(npartial), dim3(256), 0, 0, pred, target, mask, grad, partial,
                       pixels, cu, ctot, inv_count);
    return status();
}

int mplbws_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr, float b1,
                float b2, float eps, int step) {
    if (!p || !g || !m || !v || count < 0 || step < 1) return MPLBWS_ERR_ARG;
    const float c1 = 1.f / (1.f - powf(b1, (float)step));
    const float c2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(count)), dim3(256), 0, 0, p, g, m, v, count, lr, b1,
                       b2, eps, c1, c2);
    return status();
}

int mplbws_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : MPLBWS_ERR_HIP; }

int mplbws_alloc(void** p, int64_t bytes) {
    if (!p || bytes < 0) return MPLBWS_ERR_ARG;
    *p = nullptr;
    if (bytes == 0) return 0;
    return hipMalloc(p, (size_t)bytes) == hipSuccess ? 0 : MPLBWS_ERR_HIP;
}

int mplbws_free(void* p) {
    if (!p) return 0;
    return hipFree(p) == hipSuccess ? 0 : MPLBWS_ERR_HIP;
}

int mplbws_upload(void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return MPLBWS_ERR_ARG;
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess
               ? 0
               : MPLBWS_ERR_HIP;
}

int mplbws_download(void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return MPLBWS_ERR_ARG;
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess
               ? 0
               : MPLBWS_ERR_HIP;
}

int mplbws_zero(void* p, int64_t bytes) {
    if (!p || bytes < 0 || bytes % 4) return MPLBWS_ERR_ARG;
    hipLaunchKernelGGL(k_zero, dim3(grid_for(bytes / 4)), dim3(256), 0, 0,
                       static_cast<float*>(p), bytes / 4);
    return status();
}

}  // extern "C"
