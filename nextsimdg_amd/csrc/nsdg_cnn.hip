// The warm-start network: 3x3 convolutions as implicit GEMMs on MFMA
// (v_mfma_f32_16x16x32_bf16: A 16x32, B 32x16, D 16x16 FP32), with
//   A[m][k]: lane l holds m = l % 16, k = 8 (l / 16) + i, i = 0..7
//   B[k][n]: lane l holds n = l % 16, k = 8 (l / 16) + i
//   D[m][n]: lane l holds n = l % 16, m = 4 (l / 16) + i, i = 0..3.
// Forward and backward-data: m = 16 consecutive pixels of a row, k = the 32 input
// channels of one tap, n = 16 output channels; a wave holds the 9 x 2 weight fragments
// of a layer in registers and walks over pixel tiles. Weight gradient: m = input
// channel, n = output channel, k = pixels, staged channel-major in LDS.
#include "nsdg_cnn.hpp"

#include <cmath>

namespace nsdg {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t toBF16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float fromBF16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ bf16x8 frag(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p)); }

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// BF16 weight fragments from the FP32 master values W[l][co][ci][tap]:
//   wf[l][tap][j][lane][i] = W[co = 16 j + lane % 16][ci = 8 (lane / 16) + i][tap]
//   wb[l][tap][j][lane][i] = W[co = 8 (lane / 16) + i][ci = 16 j + lane % 16][8 - tap]
// (wb: the transposed, flipped kernel of the data gradient).
__global__ void k_pack(const float* params, uint16_t* wf, uint16_t* wb)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= LAYERS * FRAG)
        return;
    const int l = idx / FRAG, r = idx % FRAG;
    const int i = r % 8, lane = (r / 8) % 64, j = (r / 512) % 2, tap = r / 1024;
    const float* W = params + static_cast<size_t>(l) * PSZ;
    {
        const int co = 16 * j + lane % 16, ci = 8 * (lane / 16) + i;
        wf[idx] = toBF16(W[(co * C + ci) * TAPS + tap]);
    }
    {
        const int co = 8 * (lane / 16) + i, ci = 16 * j + lane % 16;
        wb[idx] = toBF16(W[(co * C + ci) * TAPS + (TAPS - 1 - tap)]);
    }
}

enum Epi { EPI_RELU = 0, EPI_OUT = 1, EPI_MASK = 2 };

// out = epilogue(sum_tap in(p + offset(tap)) W(tap)), 'same' zero padding.
//   EPI_RELU: out16[p][co] = bf16(max(acc + bias, 0))          (32 channels)
//   EPI_OUT:  out32[p][co] = acc + bias, co < 8                 (first 16 channels computed)
//   EPI_MASK: out16[p][co] = bf16(acc) where mask[p][co] > 0, else 0
// A block's band is 64 pixels (16 per wave) of ROWS consecutive output rows. Its input,
// rows y0 - 1 .. y0 + ROWS and columns x0 - 1 .. x0 + 64 with zeros outside the image,
// is staged in LDS once (16-byte chunks, coalesced), at a pixel stride of 80 bytes so that
// the 16 pixels of a fragment read fall in distinct banks. Input row y0 + r is read once
// per dx and feeds the output rows y0 + r - dy of its three taps; each output still sums
// its taps in tap order (dy, then dx). The outputs go back through LDS and leave as
// 16-byte stores of whole pixels. The next band's input is loaded into registers while
// this one computes.
constexpr int kConvRows = 8, kBandX = 16 * kWaves, kInRows = kConvRows + 2, kInCols = kBandX + 2;
constexpr int kPix = 80; // bytes per staged input pixel (64 used)
constexpr int kConvLds = kInRows * kInCols * kPix;
static_assert(kConvRows * kBandX * C * 2 <= kConvLds, "the output stage fits in the input stage");
template <int EPI>
__global__ __launch_bounds__(kThreads) void k_conv(const uint16_t* __restrict__ in, const uint16_t* __restrict__ w,
    const float* __restrict__ bias, const uint16_t* __restrict__ mask, uint16_t* __restrict__ out16,
    float* __restrict__ out32, int n, int ny, int nx)
{
    constexpr int NT = EPI == EPI_OUT ? 1 : 2, ROWS = kConvRows;
    __shared__ __align__(16) unsigned char lds[kConvLds];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    bf16x8 wr[TAPS][NT];
#pragma unroll
    for (int t = 0; t < TAPS; ++t)
#pragma unroll
        for (int j = 0; j < NT; ++j)
            wr[t][j] = frag(w + ((t * 2 + j) * 64 + lane) * 8);
    const int bandsX = (nx + kBandX - 1) / kBandX, bandsY = (ny + ROWS - 1) / ROWS;
    const long nbands = static_cast<long>(n) * bandsY * bandsX;
    // The input chunks of a band, loaded into registers one band ahead: thread t holds
    // chunks t, t + kThreads, ... (4 chunks of 16 bytes per pixel).
    constexpr int NC = (kInRows * kInCols * 4 + kThreads - 1) / kThreads;
    uint4 pre[NC];
    const auto load = [&](long band) {
        const int bx0 = static_cast<int>(band % bandsX) * kBandX;
        const long rb = band / bandsX;
        const int by0 = static_cast<int>(rb % bandsY) * ROWS;
        const long bimg = rb / bandsY;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int c = threadIdx.x + k * kThreads;
            const int q = c % 4, col = (c / 4) % kInCols, row = c / (4 * kInCols);
            const int xp = bx0 - 1 + col, yp = by0 - 1 + row;
            pre[k] = make_uint4(0, 0, 0, 0);
            if (c < kInRows * kInCols * 4 && xp >= 0 && xp < nx && yp >= 0 && yp < ny)
                pre[k] = *reinterpret_cast<const uint4*>(in + ((bimg * ny + yp) * nx + xp) * C + 8 * q);
        }
    };
    if (blockIdx.x < nbands)
        load(blockIdx.x);
    for (long band = blockIdx.x; band < nbands; band += gridDim.x) {
        const int x0 = static_cast<int>(band % bandsX) * kBandX;
        const long rowBand = band / bandsX; // img * bandsY + by
        const int y0 = static_cast<int>(rowBand % bandsY) * ROWS;
        const long img = rowBand / bandsY;
        // stage the input band, then start loading the next one
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int c = threadIdx.x + k * kThreads;
            if (c < kInRows * kInCols * 4) {
                const int q = c % 4, col = (c / 4) % kInCols, row = c / (4 * kInCols);
                *reinterpret_cast<uint4*>(lds + (row * kInCols + col) * kPix + 16 * q) = pre[k];
            }
        }
        __syncthreads();
        if (band + gridDim.x < nbands)
            load(band + gridDim.x);
        floatx4 acc[ROWS][NT];
#pragma unroll
        for (int i = 0; i < ROWS; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j)
                acc[i][j] = floatx4 { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int r = -1; r <= ROWS; ++r) {
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int col = 16 * wv + lane % 16 + dx + 1;
                const bf16x8 a = __builtin_bit_cast(bf16x8,
                    *reinterpret_cast<const uint4*>(lds + ((r + 1) * kInCols + col) * kPix + 16 * (lane / 16)));
#pragma unroll
                for (int dy = 1; dy >= -1; --dy) { // output row y0 + r - dy, tap (dy, dx)
                    const int i = r - dy;
                    if (i < 0 || i >= ROWS)
                        continue;
                    const int t = (dy + 1) * 3 + (dx + 1);
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wr[t][j], acc[i][j], 0, 0, 0);
                }
            }
        }
        __syncthreads(); // every fragment read before the outputs overwrite the stage
        // D[m][n]: lane holds channel n = lane % 16 (+ 16 j) of pixels m = 4 (lane / 16) + i
        if constexpr (EPI == EPI_OUT) {
            float* st = reinterpret_cast<float*>(lds); // [ROWS][kBandX][OUT]
            const int co = lane % 16;
            if (co < OUT) {
#pragma unroll
                for (int ir = 0; ir < ROWS; ++ir)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        st[(ir * kBandX + 16 * wv + 4 * (lane / 16) + i) * OUT + co] = acc[ir][0][i] + bias[co];
            }
        } else {
            uint16_t* st = reinterpret_cast<uint16_t*>(lds); // [ROWS][kBandX][C]
#pragma unroll
            for (int ir = 0; ir < ROWS; ++ir)
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    const int co = 16 * j + lane % 16;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float val = acc[ir][j][i];
                        st[(ir * kBandX + 16 * wv + 4 * (lane / 16) + i) * C + co]
                            = EPI == EPI_RELU ? toBF16(fmaxf(val + bias[co], 0.0f)) : toBF16(val);
                    }
                }
        }
        __syncthreads();
        // whole pixels out: 16-byte chunks (4 per BF16 pixel, 2 per FP32 output pixel)
        constexpr int CH = EPI == EPI_OUT ? 2 : 4;
        for (int c = threadIdx.x; c < ROWS * kBandX * CH; c += kThreads) {
            const int q = c % CH, px = (c / CH) % kBandX, ir = c / (CH * kBandX);
            const int x = x0 + px, y = y0 + ir;
            if (x >= nx || y >= ny)
                continue;
            const long p = (img * ny + y) * nx + x;
            const uint4 v = *reinterpret_cast<const uint4*>(lds + (static_cast<size_t>(ir * kBandX + px) * CH + q) * 16);
            if constexpr (EPI == EPI_OUT) {
                *reinterpret_cast<uint4*>(out32 + p * OUT + 4 * q) = v;
            } else if constexpr (EPI == EPI_RELU) {
                *reinterpret_cast<uint4*>(out16 + p * C + 8 * q) = v;
            } else { // keep the gradient where the layer's ReLU was active
                const uint4 mk = *reinterpret_cast<const uint4*>(mask + p * C + 8 * q);
                const uint32_t vv[4] = { v.x, v.y, v.z, v.w }, mm[4] = { mk.x, mk.y, mk.z, mk.w };
                uint32_t o[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t lo = fromBF16(static_cast<uint16_t>(mm[k] & 0xffffu)) > 0.0f ? (vv[k] & 0xffffu) : 0u;
                    const uint32_t hi = fromBF16(static_cast<uint16_t>(mm[k] >> 16)) > 0.0f ? (vv[k] & 0xffff0000u) : 0u;
                    o[k] = lo | hi;
                }
                *reinterpret_cast<uint4*>(out16 + p * C + 8 * q) = make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
        __syncthreads(); // the stage is free for the next band
    }
}

// Loss sum_p w_p sum_{c<8} (z - t)^2 / (8 npix) and its gradient with respect to z, stored
// as BF16 over 32 channels (zero for c >= 8). One FP64 partial sum per block.
__global__ __launch_bounds__(kThreads) void k_loss(const float* z, const float* t, const float* w, long npix,
    uint16_t* dz, double* part)
{
    __shared__ double red[kWaves];
    const float scale = 2.0f / (8.0f * static_cast<float>(npix));
    double acc = 0.0;
    for (long p = static_cast<long>(blockIdx.x) * kThreads + threadIdx.x; p < npix;
         p += static_cast<long>(gridDim.x) * kThreads) {
        const float wp = w ? w[p] : 1.0f;
#pragma unroll
        for (int c = 0; c < OUT; ++c) {
            const float d = z[p * OUT + c] - t[p * OUT + c];
            acc += static_cast<double>(wp) * d * d;
            dz[p * C + c] = toBF16(scale * wp * d);
        }
#pragma unroll
        for (int c = OUT; c < C; ++c)
            dz[p * C + c] = 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x % 64 == 0)
        red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int k = 0; k < kWaves; ++k)
            sum += red[k];
        part[blockIdx.x] = sum / (8.0 * static_cast<double>(npix));
    }
}

// Weight gradient dW[co][ci][tap] = sum_p a(p + offset(tap))[ci] dz(p)[co] and bias
// gradient sum_p dz(p)[co], as per-block partials part[block][tap][ci][co] (+ [co]).
// A block takes 64-pixel row segments; wave w computes the (ci tile w / 2, co tile w % 2)
// block of every tap.
constexpr int SEG = 64;
__global__ __launch_bounds__(kThreads) void k_wgrad(const uint16_t* __restrict__ a, const uint16_t* __restrict__ dz,
    int n, int ny, int nx, float* __restrict__ part)
{
    __shared__ uint16_t at[3][C][SEG + 2];
    __shared__ __attribute__((aligned(16))) uint16_t gt[C][SEG];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int mi = wave / 2, nj = wave % 2;
    floatx4 acc[TAPS];
#pragma unroll
    for (int t = 0; t < TAPS; ++t)
        acc[t] = floatx4 { 0.f, 0.f, 0.f, 0.f };
    float db = 0.0f;
    const int segs = (nx + SEG - 1) / SEG;
    const long items = static_cast<long>(n) * ny * segs;
    for (long it = blockIdx.x; it < items; it += gridDim.x) {
        const int seg = static_cast<int>(it % segs);
        const long rowId = it / segs;
        const int y = static_cast<int>(rowId % ny);
        const long img = rowId / ny;
        const int x0 = SEG * seg;
        __syncthreads();
        for (int q = threadIdx.x; q < 3 * (SEG + 2) * 4; q += kThreads) {
            const int cg = q % 4, px = (q / 4) % (SEG + 2), row = q / (4 * (SEG + 2));
            const int xp = x0 - 1 + px, yp = y - 1 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (xp >= 0 && xp < nx && yp >= 0 && yp < ny)
                v = *reinterpret_cast<const uint4*>(a + ((img * ny + yp) * nx + xp) * C + 8 * cg);
            const uint32_t wv[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                at[row][8 * cg + 2 * k][px] = static_cast<uint16_t>(wv[k] & 0xFFFFu);
                at[row][8 * cg + 2 * k + 1][px] = static_cast<uint16_t>(wv[k] >> 16);
            }
        }
        for (int q = threadIdx.x; q < SEG * 4; q += kThreads) {
            const int cg = q % 4, px = q / 4;
            const int xp = x0 + px;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (xp < nx)
                v = *reinterpret_cast<const uint4*>(dz + (rowId * nx + xp) * C + 8 * cg);
            const uint32_t wv[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gt[8 * cg + 2 * k][px] = static_cast<uint16_t>(wv[k] & 0xFFFFu);
                gt[8 * cg + 2 * k + 1][px] = static_cast<uint16_t>(wv[k] >> 16);
            }
        }
        __syncthreads();
        if (threadIdx.x < C)
            for (int px = 0; px < SEG; ++px)
                db += fromBF16(gt[threadIdx.x][px]);
#pragma unroll
        for (int kc = 0; kc < SEG / 32; ++kc) {
            const bf16x8 b = frag(&gt[16 * nj + lane % 16][32 * kc + 8 * (lane / 16)]);
#pragma unroll
            for (int t = 0; t < TAPS; ++t) {
                const int dyi = t / 3, dxi = t % 3;
                const uint16_t* src = &at[dyi][16 * mi + lane % 16][32 * kc + 8 * (lane / 16) + dxi];
                bf16x8 av;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    av[i] = __builtin_bit_cast(__bf16, src[i]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, acc[t], 0, 0, 0);
            }
        }
    }
    float* P = part + static_cast<size_t>(blockIdx.x) * PSZ;
#pragma unroll
    for (int t = 0; t < TAPS; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int ci = 16 * mi + 4 * (lane / 16) + i, co = 16 * nj + lane % 16;
            P[(t * C + ci) * C + co] = acc[t][i];
        }
    if (threadIdx.x < C)
        P[WSZ + threadIdx.x] = db;
}
// Sum of the partials in block order, into the parameter layout [co][ci][tap], bias. The
// loads go out 16 at a time ahead of their (in-order) additions.
__global__ void k_wgrad_reduce(const float* part, int nb, float* g)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= PSZ)
        return;
    float sum = 0.0f;
    int b = 0;
    for (; b + 16 <= nb; b += 16) {
        float v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            v[k] = part[static_cast<size_t>(b + k) * PSZ + e];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            sum += v[k];
    }
    for (; b < nb; ++b)
        sum += part[static_cast<size_t>(b) * PSZ + e];
    if (e < WSZ) {
        const int t = e / (C * C), ci = (e / C) % C, co = e % C;
        g[(co * C + ci) * TAPS + t] = sum;
    } else {
        g[e] = sum;
    }
}

__global__ void k_adam(float* p, float* m, float* v, const float* g, int n, float lr, float b1, float b2, float eps,
    float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------- host side
namespace {
uint64_t splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double normalAt(uint64_t seed, uint64_t i)
{
    const double u1 = (static_cast<double>(splitmix(seed ^ (2 * i)) >> 11) + 0.5) * 0x1p-53;
    const double u2 = (static_cast<double>(splitmix(seed ^ (2 * i + 1)) >> 11) + 0.5) * 0x1p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}
unsigned gridFor(long work, long perBlock, long cap)
{
    const long b = (work + perBlock - 1) / perBlock;
    return static_cast<unsigned>(b < 1 ? 1 : (b > cap ? cap : b));
}
constexpr int kWgradBlocks = 512;
} // namespace

Net::Net(uint64_t seed)
{
    NSDG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<float> h(NPARAM, 0.0f);
    const double std0 = std::sqrt(2.0 / (C * TAPS));
    for (int l = 0; l < LAYERS; ++l)
        for (int co = 0; co < C; ++co)
            for (int k = 0; k < C * TAPS; ++k) {
                const size_t i = static_cast<size_t>(l) * PSZ + co * C * TAPS + k;
                double wv = std0 * normalAt(seed, i);
                if (l == LAYERS - 1)
                    wv = co < OUT ? 0.1 * wv : 0.0;
                h[i] = static_cast<float>(wv);
            }
    for (auto* b : { &params, &m, &v, &grad })
        b->alloc(NPARAM);
    params.upload(h.data(), NPARAM, s);
    std::vector<float> z(NPARAM, 0.0f);
    for (auto* b : { &m, &v, &grad })
        b->upload(z.data(), NPARAM, s);
    wf.alloc(static_cast<size_t>(LAYERS) * FRAG);
    wb.alloc(static_cast<size_t>(LAYERS) * FRAG);
    part.alloc(static_cast<size_t>(kWgradBlocks) * PSZ);
    lossPart.alloc(1024);
    pack();
    sync();
}
Net::~Net()
{
    if (s)
        (void)hipStreamDestroy(s);
}
void Net::sync()
{
    NSDG_HIP(hipStreamSynchronize(s));
    if (prof.on)
        prof.collect();
}
void Net::pack()
{
    k_pack<<<(LAYERS * FRAG + 255) / 256, 256, 0, s>>>(params.p, wf.p, wb.p);
    checkLaunch("cnn_pack");
}
void Net::reserve(size_t npix)
{
    if (npix <= pixels)
        return;
    for (int l = 1; l < LAYERS; ++l)
        act[l].alloc(npix * C);
    dz0.alloc(npix * C);
    dz1.alloc(npix * C);
    out.alloc(npix * OUT);
    pixels = npix;
}

namespace {
void checkShape(const void* x, int n, int ny, int nx)
{
    require(n >= 1 && ny >= 1 && nx >= 16, "network input needs n >= 1, ny >= 1, nx >= 16");
    require(nx % 16 == 0, "network input width must be a multiple of 16");
    require(x != nullptr && reinterpret_cast<uintptr_t>(x) % 16 == 0, "network input must be 16-byte aligned");
}
template <int EPI>
void conv(hipStream_t s, Profiler& prof, const char* name, const uint16_t* in, const uint16_t* w, const float* bias,
    const uint16_t* mask, uint16_t* out16, float* out32, int n, int ny, int nx)
{
    Timed t(prof, s, name);
    const long bands = static_cast<long>(n) * ((ny + kConvRows - 1) / kConvRows) * ((nx + kBandX - 1) / kBandX);
    k_conv<EPI><<<gridFor(bands, 1, 8192), kThreads, 0, s>>>(in, w, bias, mask, out16, out32, n, ny, nx);
    checkLaunch(name);
}
} // namespace

void Net::forward(const uint16_t* x, int n, int ny, int nx, float* y, bool)
{
    checkShape(x, n, ny, nx);
    const size_t npix = static_cast<size_t>(n) * ny * nx;
    reserve(npix);
    float* dst = y ? y : out.p;
    const float* b = params.p;
    conv<EPI_RELU>(s, prof, "cnn_conv", x, wf.p, b + 0 * PSZ + WSZ, nullptr, act[1].p, nullptr, n, ny, nx);
    conv<EPI_RELU>(s, prof, "cnn_conv", act[1].p, wf.p + FRAG, b + 1 * PSZ + WSZ, nullptr, act[2].p, nullptr, n, ny, nx);
    conv<EPI_RELU>(s, prof, "cnn_conv", act[2].p, wf.p + 2 * FRAG, b + 2 * PSZ + WSZ, nullptr, act[3].p, nullptr, n, ny,
        nx);
    conv<EPI_OUT>(s, prof, "cnn_conv_out", act[3].p, wf.p + 3 * FRAG, b + 3 * PSZ + WSZ, nullptr, nullptr, dst, n, ny,
        nx);
}

double Net::lossGrad(const uint16_t* x, const float* t, const float* w, int n, int ny, int nx)
{
    require(t != nullptr, "target is NULL");
    forward(x, n, ny, nx, nullptr, true);
    const long npix = static_cast<long>(n) * ny * nx;
    const unsigned nbLoss = gridFor(npix, kThreads, 1024);
    {
        Timed tm(prof, s, "cnn_loss");
        k_loss<<<nbLoss, kThreads, 0, s>>>(out.p, t, w, npix, dz0.p, lossPart.p);
        checkLaunch("cnn_loss");
    }
    const long items = static_cast<long>(n) * ny * ((nx + SEG - 1) / SEG);
    const unsigned nbW = gridFor(items, 1, kWgradBlocks);
    auto wgrad = [&](const uint16_t* a, const uint16_t* dz, int layer) {
        Timed tm(prof, s, "cnn_wgrad");
        k_wgrad<<<nbW, kThreads, 0, s>>>(a, dz, n, ny, nx, part.p);
        checkLaunch("cnn_wgrad");
        k_wgrad_reduce<<<(PSZ + 255) / 256, 256, 0, s>>>(part.p, static_cast<int>(nbW), grad.p + layer * PSZ);
        checkLaunch("cnn_wgrad_reduce");
    };
    // layer 3 (linear output): dz0 = dL/dz4
    wgrad(act[3].p, dz0.p, 3);
    conv<EPI_MASK>(s, prof, "cnn_conv_bwd", dz0.p, wb.p + 3 * FRAG, nullptr, act[3].p, dz1.p, nullptr, n, ny, nx);
    wgrad(act[2].p, dz1.p, 2);
    conv<EPI_MASK>(s, prof, "cnn_conv_bwd", dz1.p, wb.p + 2 * FRAG, nullptr, act[2].p, dz0.p, nullptr, n, ny, nx);
    wgrad(act[1].p, dz0.p, 1);
    conv<EPI_MASK>(s, prof, "cnn_conv_bwd", dz0.p, wb.p + 1 * FRAG, nullptr, act[1].p, dz1.p, nullptr, n, ny, nx);
    wgrad(x, dz1.p, 0);
    std::vector<double> lp(nbLoss);
    lossPart.download(lp.data(), nbLoss, s);
    sync();
    double loss = 0.0;
    for (double v : lp)
        loss += v;
    return loss;
}

void Net::adam(float lr)
{
    ++adamStep;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, static_cast<float>(adamStep));
    const float bc2 = 1.0f - std::pow(b2, static_cast<float>(adamStep));
    {
        Timed tm(prof, s, "cnn_adam");
        k_adam<<<(NPARAM + 255) / 256, 256, 0, s>>>(params.p, m.p, v.p, grad.p, NPARAM, lr, b1, b2, 1e-8f, bc1, bc2);
        checkLaunch("cnn_adam");
    }
    pack();
    sync();
}

} // namespace cnn
} // namespace nsdg
