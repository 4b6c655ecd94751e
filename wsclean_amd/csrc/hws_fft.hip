// LDS-staged mixed-radix Stockham FFT for the per-layer 2-D transforms
// (the reference runs fftw[f]_plan_dft_2d per layer,
// gridding/wstackinggridder.cpp:200-239, 282-321).
//
// A 2-D transform of a batch of H x W layers is two launches:
//   rows:    one workgroup per (row block, layer); a row of W complex values
//            is loaded with coalesced float2 loads, transformed through the
//            Stockham stages in LDS, and written back in place;
//   columns: one workgroup per (column block of Bc columns, layer); the Bc
//            columns are gathered into LDS as [Bc][H+1] (padded to spread
//            banks), transformed, and scattered back. Column blocks that
//            share 128-byte lines are mapped onto the same XCD so that they
//            share its L2.
// Each Stockham stage reads its R inputs into registers, synchronises, and
// writes its R outputs in natural (autosorted) order, so one LDS buffer per
// transform suffices. Radices 8, 4, 2, 3, 5, 7.
#include "hws_internal.h"

#include <cmath>

namespace hws {

namespace {

// The butterflies and twiddle products fuse only where written (fmaf), so a
// transform computes the same bits in every kernel that inlines it; left to
// itself, the compiler contracts a product into the following sum in some
// inlining contexts and not in others.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
// multiply by sign * i
template <int SIGN>
__device__ __forceinline__ float2 mulI(float2 a) {
  return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int SIGN, int R>
struct Butterfly;

template <int SIGN>
struct Butterfly<SIGN, 2> {
  __device__ __forceinline__ static void run(float2* v) {
#pragma clang fp contract(off)
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int SIGN>
struct Butterfly<SIGN, 4> {
  __device__ __forceinline__ static void run(float2* v) {
#pragma clang fp contract(off)
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 b0 = cadd(v[1], v[3]), b1 = mulI<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  }
};

template <int SIGN>
struct Butterfly<SIGN, 8> {
  __device__ __forceinline__ static void run(float2* v) {
#pragma clang fp contract(off)
    // radix-2 over even/odd, then radix-4 on each half
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Butterfly<SIGN, 4>::run(e);
    Butterfly<SIGN, 4>::run(o);
    const float h = 0.70710678118654752440f;
    // o[k] *= w8^k, w8 = exp(SIGN * 2 pi i / 8)
    o[1] = make_float2(h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x));
    o[2] = mulI<SIGN>(o[2]);
    o[3] = make_float2(h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// Odd radix: direct DFT with exp(SIGN 2 pi i k/R) constants.
template <int SIGN, int R>
struct Butterfly {
  __device__ __forceinline__ static void run(float2* v) {
#pragma clang fp contract(off)
    float2 res[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int m = (r * k) % R;
        float s, c;
        // constants folded at compile time
        c = (float)__builtin_cos(2.0 * 3.14159265358979323846 * m / R);
        s = (float)(SIGN * __builtin_sin(2.0 * 3.14159265358979323846 * m / R));
        acc = cadd(acc, cmul(v[r], make_float2(c, s)));
      }
      res[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = res[k];
  }
};

constexpr int kMaxValues = 16;  // complex values held per thread per stage

// LDS position of element e of a line of the runtime-radix transforms: one
// pad value after every 16, so that the first stage's writes (stride R = 16,
// 8, 4 or 2 values across the lanes) fall in distinct banks.
__device__ __forceinline__ int lp(int e) { return e + (e >> 4); }

// Line pitch (in values) of those transforms for length n: the padded length
// plus one.
__host__ __device__ constexpr int ldsPitch(int n) { return n + (n >> 4) + 1; }

// Division by a divisor that is the same on every lane: x / d ==
// umulhi(x, ceil(2^32 / d)) for d >= 2 whenever x * d < 2^32 (here x and d
// stay below 2^16).
struct UniformDiv {
  uint32_t m;
  int d;
};
__device__ __forceinline__ UniformDiv uniformDiv(int d) {
  return UniformDiv{d > 1 ? 0xFFFFFFFFu / uint32_t(d) + 1u : 0u, d};
}
__device__ __forceinline__ int divBy(int x, UniformDiv u) {
  return u.d == 1 ? x : int(__umulhi(uint32_t(x), u.m));
}

// One Stockham stage of radix R over nfft transforms of length n stored at
// buf + f * pitch (element e at lp(e)). Butterfly b of transform f reads
// x[b + r*B] (B = n/R), applies twiddle w^(jm*r) with jm = b % ns, and writes
// y[(b/ns)*ns*R + jm + r*ns].
template <int SIGN, int R>
__device__ __forceinline__ void stockhamStage(float2* buf, int n, int pitch,
                                              int nfft, int ns,
                                              const float2* __restrict__ tw,
                                              int tid, int nthr) {
  constexpr int kBatches = kMaxValues / R;
  // The batch indices below depend only on (tid, n, R, nthr), so the compiler
  // would hoist those of every radix out of the stage loop of fftLds and keep
  // them all live (208 VGPRs, spills at 1024 threads); an opaque tid makes
  // each stage compute its own.
  asm volatile("" : "+v"(tid));
  const int B = n / R;
  const int total = B * nfft;
  const int twStep = n / (ns * R);
  const UniformDiv byB = uniformDiv(B), byNs = uniformDiv(ns);
  float2 v[kBatches][R];
#pragma unroll
  for (int q = 0; q < kBatches; ++q) {
    const int t = tid + q * nthr;
    if (t < total) {
      const int f = divBy(t, byB), b = t - f * B;
      const int jm = b - divBy(b, byNs) * ns;
      const float2* x = buf + f * pitch;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 val = x[lp(b + r * B)];
        if (r != 0 && jm != 0) val = cmul(val, tw[jm * r * twStep]);
        v[q][r] = val;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kBatches; ++q) {
    const int t = tid + q * nthr;
    if (t < total) {
      const int f = divBy(t, byB), b = t - f * B;
      Butterfly<SIGN, R>::run(v[q]);
      const int bq = divBy(b, byNs), jm = b - bq * ns;
      const int idxD = bq * ns * R + jm;
      float2* y = buf + f * pitch;
#pragma unroll
      for (int r = 0; r < R; ++r) y[lp(idxD + r * ns)] = v[q][r];
    }
  }
  __syncthreads();
}

// Stage radices are packed 4 bits per stage (stage 0 lowest) so that the
// per-stage dispatch reads a scalar rather than an indexed argument array.
// Lines are stored padded (element e of line f at buf[f * pitch + lp(e)],
// pitch >= ldsPitch(n)).
template <int SIGN>
__device__ __forceinline__ void fftLds(float2* buf, int pitch, int nfft, int n,
                       unsigned long long radices, const float2* tw, int tid,
                       int nthr) {
  int ns = 1;
  while (radices != 0ull) {
    const int code = int(radices & 15ull);
    const int R = code == 1 ? 16 : code;  // (radix 16 is packed as 1)
    radices >>= 4;
    switch (R) {
      case 16: stockhamStage<SIGN, 16>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 8: stockhamStage<SIGN, 8>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 4: stockhamStage<SIGN, 4>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 2: stockhamStage<SIGN, 2>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 3: stockhamStage<SIGN, 3>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 5: stockhamStage<SIGN, 5>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      case 7: stockhamStage<SIGN, 7>(buf, n, pitch, nfft, ns, tw, tid, nthr); break;
      default: break;
    }
    ns *= R;
  }
}

// ---- power-of-two lengths: every index is a compile-time shift or mask.
// Stage radices are 8 while the remaining length allows, then 4, then 2.
template <int SIGN, int R, int N, int NS, int NT, int NFFT>
__device__ __forceinline__ void pow2Stage(float2* buf, int pitch, int nfft,
                                          const float2* __restrict__ tw, int tid) {
  constexpr int B = N / R;
  constexpr int TOTAL = B * NFFT;
  constexpr int BATCHES = (TOTAL + NT - 1) / NT;
  constexpr int TWSTEP = N / (NS * R);
  float2 v[BATCHES][R];
#pragma unroll
  for (int q = 0; q < BATCHES; ++q) {
    const int t = tid + q * NT;
    const int f = t / B, b = t % B;
    if (f < nfft) {
      const int jm = b % NS;
      const float2* x = buf + f * pitch;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 val = x[b + r * B];
        if (NS > 1 && r != 0) val = cmul(val, tw[jm * r * TWSTEP]);
        v[q][r] = val;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BATCHES; ++q) {
    const int t = tid + q * NT;
    const int f = t / B, b = t % B;
    if (f < nfft) {
      Butterfly<SIGN, R>::run(v[q]);
      const int jm = b % NS;
      float2* y = buf + f * pitch + (b / NS) * NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) y[r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

template <int SIGN, int N, int NT, int NFFT, int NS = 1>
__device__ __forceinline__ void pow2Fft(float2* buf, int pitch, int nfft,
                                        const float2* __restrict__ tw, int tid) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (REM % 8 == 0) ? 8 : (REM % 4 == 0) ? 4 : 2;
    pow2Stage<SIGN, R, N, NS, NT, NFFT>(buf, pitch, nfft, tw, tid);
    pow2Fft<SIGN, N, NT, NFFT, NS * R>(buf, pitch, nfft, tw, tid);
  }
}

// ---- blocked layer layout: element (x, y) of a W x H layer is stored at
// ((y >> 2) * W + x) * 4 + (y & 3), i.e. groups of four rows interleaved.
// A row block (four rows) is then one contiguous run of 4W values and the
// four columns x0..x0+3 form 128-byte segments, so both passes of the 2-D
// transform read and write HBM fully coalesced. Each pass keeps four
// transforms of length N (N/4 threads, 16 values per thread) in flight:
// stage 0 reads HBM into registers, stages exchange through one padded LDS
// buffer, the last stage writes HBM. Radix 16 stages, then 8, 4 or 2.

__device__ __forceinline__ int padIdx(int e) { return e + (e >> 6); }

template <int SIGN>
struct Butterfly<SIGN, 16> {
  __device__ __forceinline__ static void run(float2* v) {
#pragma clang fp contract(off)
    // 4 x 4: DFT4 over stride-4 inputs, twiddle w16^(n1 k2), DFT4 again
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[n2 + 4 * n1];
      Butterfly<SIGN, 4>::run(a[n2]);  // a[n2][k1]
    }
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    // twiddle a[n2][k1] *= w16^(n2 k1), w16 = exp(SIGN 2 pi i / 16)
    const float2 w1 = make_float2(c1, SIGN * s1), w2 = make_float2(h, SIGN * h),
                 w3 = make_float2(s1, SIGN * c1), w4 = make_float2(0.0f, float(SIGN)),
                 w6 = make_float2(-h, SIGN * h), w9 = make_float2(-c1, -SIGN * s1);
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = cmul(a[2][2], w4);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Butterfly<SIGN, 4>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

template <bool ROWS>
__device__ __forceinline__ size_t blockedAddr(int pos, int f, int W) {
  // ROWS: pos = x, f = row in block; COLS: pos = y, f = column in block
  if (ROWS) return size_t(pos) * 4 + f;
  return (size_t(pos >> 2) * W + f) * 4 + (pos & 3);
}

// live (may be null): one flag per tile-sized stretch of the block's rows
// (rows pass: per uv tile along x) or columns (columns pass: per tile row
// along y), indexed by pos >> tshift (tshift = log2 of the tile edge).
// liveIn: cells of a stretch that is not live read as zero (their memory is
// not read, and need not have been written); otherwise they are not written.
// The layer cell (x, y) of a forward transform computed from the model image
// (the arithmetic of imageToLayersKernel, so the results are identical). REF:
// the reference precision (float n', FP64 sincos).
template <bool REF>
__device__ __forceinline__ float2 modelCell(const ModelSource& ms, double w, int x, int y, int W,
                                           int H) {
  int yDest = y + H / 2;
  if (yDest >= H) yDest -= H;
  int xDest = (W - x) + W / 2;
  if (xDest >= W) xDest -= W;
  const double m = pixelM(yDest, H, ms.dm, ms.mShift);
  const double l = pixelL(xDest, W, ms.dl, ms.lShift);
  double np = nPrimeSeries(l, m);
  if (REF) np = double(float(np));
  const size_t src = size_t(yDest) * W + xDest;
  const float re = ms.re[src];
  const float imv = ms.cplx ? -ms.im[src] : 0.0f;
  return modelTerm<REF>(re, imv, ms.cplx != 0, w, np);
}

// Layers per workgroup of the model-fed rows pass: the workgroup reads the
// model rows of its row block once from memory and then from the L2 (the
// reference precision's FP64 sincos leaves no registers for that: one).
template <bool REF>
constexpr int modelLayers() {
  return REF ? 1 : 4;
}

// A store the pass does not read back: marked non-temporal, so the stream of
// layer writes does not evict what the caches hold (the model image the
// forward rows pass reads for every layer).
__device__ __forceinline__ void streamStore(float2* p, float2 v) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 x = {v.x, v.y};
  __builtin_nontemporal_store(x, reinterpret_cast<f2*>(p));
}

// A load of a layer element the pass reads once (non-temporal).
__device__ __forceinline__ float2 streamLoad(const float2* p) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 v = __builtin_nontemporal_load(reinterpret_cast<const f2*>(p));
  return make_float2(v.x, v.y);
}

// MODEL (rows pass of a forward transform): the first stage computes its
// inputs from the model image (ms, w; y0 = the row block's first row)
// instead of reading the layer.
template <int SIGN, int R, int N, int NS, bool ROWS, int MODEL = 0>
__device__ __forceinline__ void blockedStage(float2* lds, float2* __restrict__ g, int W,
                                             const float2* __restrict__ tw, int tid,
                                             const unsigned char* live, int tshift,
                                             bool liveIn,
                                             const ModelSource* ms = nullptr, double w = 0.0,
                                             int y0 = 0, int H = 0) {
  constexpr int B = N / R;
  constexpr int T = N / 4;
  constexpr int PER = 16 / R;
  constexpr int TWSTEP = N / (NS * R);
  constexpr bool FIRST = NS == 1;
  constexpr bool LAST = NS * R == N;
  float2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int bt = tid + q * T;
    const int f = bt & 3, b = bt >> 2;
    const int jm = b % NS;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int pos = b + r * B;
      float2 x;
      if (FIRST && MODEL)
        x = modelCell<MODEL == 2>(*ms, w, pos, y0 + f, W, H);
      else if (FIRST)
        // (the forward columns pass reads the rows pass's output once: a
        // non-temporal load; measured slower in the backward passes, whose
        // inputs the caches partly hold)
        x = (live && liveIn && !live[pos >> tshift])
                ? make_float2(0.0f, 0.0f)
                : (SIGN < 0 && !ROWS ? streamLoad(&g[blockedAddr<ROWS>(pos, f, W)])
                                     : g[blockedAddr<ROWS>(pos, f, W)]);
      else
        x = lds[padIdx(pos * 4 + f)];
      if (!FIRST && r != 0) x = cmul(x, tw[jm * r * TWSTEP]);
      v[q][r] = x;
    }
  }
  if (!FIRST) __syncthreads();  // every read of this stage precedes its writes
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int bt = tid + q * T;
    const int f = bt & 3, b = bt >> 2;
    const int jm = b % NS;
    Butterfly<SIGN, R>::run(v[q]);
    const int base = (b / NS) * NS * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int pos = base + r * NS;
      if (LAST) {
        // columns pass of a forward transform: tiles nobody reads are not written
        if (live && !liveIn && !live[pos >> tshift]) continue;
        streamStore(&g[blockedAddr<ROWS>(pos, f, W)], v[q][r]);
      }
      else
        lds[padIdx(pos * 4 + f)] = v[q][r];
    }
  }
  if (!LAST) __syncthreads();
}

template <int SIGN, int N, bool ROWS, int NS = 1, int MODEL = 0>
__device__ __forceinline__ void blockedFft(float2* lds, float2* __restrict__ g, int W,
                                           const float2* __restrict__ tw, int tid,
                                           const unsigned char* live = nullptr,
                                           int tshift = 0, bool liveIn = true,
                                           const ModelSource* ms = nullptr,
                                           double w = 0.0, int y0 = 0, int H = 0) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (REM % 16 == 0) ? 16 : REM;  // 16s, then the remaining 8, 4 or 2
    static_assert(R == 16 || R == 8 || R == 4 || R == 2, "radix");
    blockedStage<SIGN, R, N, NS, ROWS, NS == 1 ? MODEL : 0>(lds, g, W, tw, tid, live, tshift,
                                                         liveIn, ms, w, y0, H);
    blockedFft<SIGN, N, ROWS, NS * R, MODEL>(lds, g, W, tw, tid, live, tshift, liveIn, ms, w, y0,
                                             H);
  }
}

// grid (H / 4, nbatch): block = row block yb of layer blockIdx.y
template <int SIGN, int N>
__global__ void __launch_bounds__(N / 4)
fftRowsBlockedKernel(float2* __restrict__ data, int H, const float2* __restrict__ tw) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
  float2* g = data + size_t(blockIdx.y) * size_t(N) * size_t(H) + size_t(blockIdx.x) * 4 * N;
  blockedFft<SIGN, N, true>(lds, g, N, tw, threadIdx.x);
}

// Rows pass of a forward transform whose layers are computed from the model
// image on the fly (the layer memory is written by this pass only): grid
// (H / 4, nbatch), batch entry b = layer ms.layer0 + b.
template <int SIGN, int N, bool REF>
__global__ void __launch_bounds__(N / 4)
fftRowsFromModelKernel(float2* __restrict__ data, int H, int nbatch,
                       const float2* __restrict__ tw, ModelSource ms) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
#pragma unroll 1
  for (int li = 0; li < modelLayers<REF>(); ++li) {
    const int bi = int(blockIdx.y) * modelLayers<REF>() + li;
    if (bi >= nbatch) break;  // workgroup-uniform
    // the empty asm gives the model pointers, pixel scale, twiddle table and
    // thread index new values each iteration: nothing derived from them is
    // held across the layers (the registers of one layer's transform are
    // all the workgroup size allows)
    ModelSource msl = ms;
    const float2* twl = tw;
    int tid = int(threadIdx.x);
    asm volatile("" : "+s"(msl.re), "+s"(msl.im), "+s"(msl.dl), "+s"(twl), "+v"(tid));
    float2* g = data + size_t(bi) * size_t(N) * size_t(H) + size_t(blockIdx.x) * 4 * N;
    const double w = layerToW(ms.map, ms.layer0 + bi);
    blockedFft<SIGN, N, true, 1, REF ? 2 : 1>(lds, g, N, twl, tid, nullptr, 0, true, &msl, w,
                                              int(blockIdx.x) * 4, H);
  }
}

// The same over layers whose uv tiles without samples were never written:
// occ holds one bit per (layer, tile row, tile) bin of the layer set
// (bin = (layer * tilesY + ty) * tilesX + tx), layer0 is the layer of batch
// entry 0, the tile edge is 1 << tshift (>= 4, so the four rows of a row
// block share one tile row).
template <int SIGN, int N>
__global__ void __launch_bounds__(N / 4)
fftRowsBlockedSparseKernel(float2* __restrict__ data, int H, const float2* __restrict__ tw,
                           const uint32_t* __restrict__ occ, int layer0, int tilesY,
                           int tshift) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
  __shared__ unsigned char live[N / 4];
  const int tilesX = N >> tshift;
  const int ty = (int(blockIdx.x) * 4) >> tshift;
  const size_t base = (size_t(layer0 + int(blockIdx.y)) * tilesY + ty) * tilesX;
  int any = 0;
  for (int t = threadIdx.x; t < tilesX; t += blockDim.x) {
    const size_t b = base + t;
    live[t] = (occ[b >> 5] >> (b & 31)) & 1u;
    any |= live[t];
  }
  float2* g = data + size_t(blockIdx.y) * size_t(N) * size_t(H) + size_t(blockIdx.x) * 4 * N;
  // no tile of the row block holds samples: its transform is zero, which the
  // columns pass knows from the same bits (it is not written)
  if (!__syncthreads_or(any)) return;
  blockedFft<SIGN, N, true>(lds, g, N, tw, threadIdx.x, live, tshift, true);
}

// Columns pass of a backward transform after fftRowsBlockedSparseKernel:
// rows of a tile row without samples were not written and read as zero.
// rowLive holds one bit per (layer, tile row) (bit = layer * tilesY + ty).
template <int SIGN, int N>
__global__ void __launch_bounds__(N / 4)
fftColsBlockedRowSparseKernel(float2* __restrict__ data, int W, const float2* __restrict__ tw,
                              const uint32_t* __restrict__ rowLive, int layer0, int tilesY,
                              int tshift) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
  __shared__ unsigned char live[N / 4];
  const int nb = gridDim.x;
  int cb = blockIdx.x;
  if ((nb & 7) == 0) cb = (cb & 7) * (nb >> 3) + (cb >> 3);
  const size_t base = size_t(layer0 + int(blockIdx.y)) * tilesY;
  for (int t = threadIdx.x; t < tilesY; t += blockDim.x) {
    const size_t b = base + t;
    live[t] = (rowLive[b >> 5] >> (b & 31)) & 1u;
  }
  __syncthreads();
  float2* g = data + size_t(blockIdx.y) * size_t(W) * size_t(N) + size_t(cb) * 16;
  blockedFft<SIGN, N, false>(lds, g, W, tw, threadIdx.x, live, tshift, true);
}

// Columns pass of a forward transform whose output is read only in some uv
// tiles (the degridding subgrids): need holds one bit per (layer, tile) bin
// as occ does for fftRowsBlockedSparseKernel; other tiles are not written.
template <int SIGN, int N>
__global__ void __launch_bounds__(N / 4)
fftColsBlockedSparseKernel(float2* __restrict__ data, int W, const float2* __restrict__ tw,
                           const uint32_t* __restrict__ need, int layer0, int tilesY,
                           int tshift) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
  __shared__ unsigned char live[N / 4];
  // column blocks in their natural order, so consecutive workgroups (dealt
  // round-robin to the XCDs) spread over all XCDs: the column blocks with
  // needed tiles lie near the uv origin, at both ends of the row, and a
  // contiguous range per XCD would leave most XCDs without work
  const int cb = blockIdx.x;
  const int tilesX = W >> tshift;
  const int tx = (cb * 4) >> tshift;
  const size_t layerBase = size_t(layer0 + int(blockIdx.y)) * tilesY;
  int any = 0;
  for (int t = threadIdx.x; t < tilesY; t += blockDim.x) {
    const size_t b = (layerBase + t) * tilesX + tx;
    live[t] = (need[b >> 5] >> (b & 31)) & 1u;
    any |= live[t];
  }
  // no tile of the column block is read afterwards: nothing to do
  if (!__syncthreads_or(any)) return;
  float2* g = data + size_t(blockIdx.y) * size_t(W) * size_t(N) + size_t(cb) * 16;
  blockedFft<SIGN, N, false>(lds, g, W, tw, threadIdx.x, live, tshift, false);
}

// grid (W / 4, nbatch): block = columns 4*cb .. 4*cb+3 of layer blockIdx.y;
// column blocks sharing an XCD are contiguous
template <int SIGN, int N>
__global__ void __launch_bounds__(N / 4)
fftColsBlockedKernel(float2* __restrict__ data, int W, const float2* __restrict__ tw) {
  __shared__ float2 lds[4 * N + (4 * N) / 64];
  const int nb = gridDim.x;
  int cb = blockIdx.x;
  if ((nb & 7) == 0) cb = (cb & 7) * (nb >> 3) + (cb >> 3);
  float2* g = data + size_t(blockIdx.y) * size_t(W) * size_t(N) + size_t(cb) * 16;
  blockedFft<SIGN, N, false>(lds, g, W, tw, threadIdx.x);
}

// ---- the backward columns pass fused with the w-projection
// One Stockham stage of radix R over NF interleaved transforms of length N
// (NF N / 16 threads, 16 values per thread), as blockedStage but with
// registers for memory: the first stage takes its inputs from io
// (io[q R + r] = element b + r N / R of transform f, for bt = tid + q T,
// f = bt mod NF, b = bt / NF) and the last stage leaves its outputs there
// (io[q R + r] = element ((b / NS) NS R + b mod NS) + r NS of transform f).
template <int SIGN, int R, int N, int NS, int NF>
__device__ __forceinline__ void regStage(float2* lds, const float2* __restrict__ tw, int tid,
                                         const float2* in, float2* out) {
  constexpr int B = N / R;
  constexpr int T = NF * N / 16;
  constexpr int PER = 16 / R;
  constexpr int TWSTEP = N / (NS * R);
  constexpr bool FIRST = NS == 1;
  constexpr bool LAST = NS * R == N;
  float2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int bt = tid + q * T;
    const int f = bt % NF, b = bt / NF;
    const int jm = b % NS;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = FIRST ? in[q * R + r] : lds[padIdx((b + r * B) * NF + f)];
      if (!FIRST && r != 0) x = cmul(x, tw[jm * r * TWSTEP]);
      v[q][r] = x;
    }
  }
  // every read of this stage precedes its writes (and, in the first stage,
  // the previous transform's last reads precede this one's writes: that
  // stage synchronised between its reads and its outputs)
  if (!FIRST) __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int bt = tid + q * T;
    const int f = bt % NF, b = bt / NF;
    const int jm = b % NS;
    Butterfly<SIGN, R>::run(v[q]);
    const int base = (b / NS) * NS * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (LAST)
        out[q * R + r] = v[q][r];
      else
        lds[padIdx((base + r * NS) * NF + f)] = v[q][r];
    }
  }
  if (!LAST) __syncthreads();
}

// The transform of in (consumed by the first stage) into out; afterFirst()
// runs once the first stage has read in (in may then be overwritten, e.g. by
// the loads of the next transform's inputs).
template <int SIGN, int N, int NF, int NS = 1, class F>
__device__ __forceinline__ void regFft(float2* lds, const float2* __restrict__ tw, int tid,
                                       const float2* in, float2* out, F&& afterFirst) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (REM % 16 == 0) ? 16 : REM;
    static_assert(R == 16 || R == 8 || R == 4 || R == 2, "radix");
    regStage<SIGN, R, N, NS, NF>(lds, tw, tid, in, out);
    if constexpr (NS == 1) afterFirst();
    regFft<SIGN, N, NF, NS * R>(lds, tw, tid, in, out, afterFirst);
  }
}

// The last stage of regFft over length N: its radix and stride.
template <int N>
struct LastStage {
  static constexpr int p16() {
    int p = 1;
    while ((N / p) % 16 == 0) p *= 16;
    return p;
  }
  static constexpr int R = N / p16() == 1 ? 16 : N / p16();
  static constexpr int NS = N / R;
};

constexpr int kProjectColumns = 2;  // columns per workgroup of the fused pass

// Buffer-resource word 3 for raw (untyped) buffer loads on CDNA.
constexpr int kRawBufferFlags = 0x00020000;
// An offset past every buffer: such a buffer load returns zero without a
// memory access.
constexpr uint32_t kOffsetOutOfRange = 0xFFFFFFF0u;

// Backward columns pass (after fftRowsBlockedSparseKernel) fused with
// projectOnImageAndCorrect (gridding/wstackinggridder.cpp:874-907) for every
// layer of the pass: a workgroup owns two columns of one image and transforms
// them layer after layer, in the pass's layer order; each transformed cell is
// rotated by its layer's w term and added (Kahan) to its pixel's sums, which
// stay in registers. So the transformed layers are never written, and the
// accumulators are read (unless the pass is the first) and written once per
// pass. Cells of tile rows without samples read as zero, as in
// fftColsBlockedRowSparseKernel. The next layer's inputs are loaded while the
// current layer is transformed. Every term is computed and added as
// projectKernel computes and adds it, so the sums are the same bits.
// grid (W / 2, images of the kind); dynamic LDS: the transform buffer and
// the row bits of the pass's layers.
template <int N, bool REF, bool CPLX>
__global__ void __launch_bounds__(kProjectColumns * N / 16)
fftColsProjectKernel(ColsProjectArgs a, const float2* __restrict__ layers,
                     const float2* __restrict__ tw, float* __restrict__ sum,
                     float* __restrict__ comp, float* __restrict__ isum,
                     float* __restrict__ icomp) {
  constexpr int NF = kProjectColumns;
  constexpr int T = NF * N / 16;
  constexpr int B0 = N / 16;  // the first stage is radix 16: one butterfly per thread
  constexpr int RL = LastStage<N>::R, NSL = LastStage<N>::NS, PERL = 16 / RL;
  extern __shared__ float2 fusedLds[];
  float2* lds = fusedLds;
  // n' of this thread's 16 output cells: [k][thread] (registers would not
  // hold them beside the transform, the prefetch and the sums)
  double* npl = reinterpret_cast<double*>(fusedLds + NF * N + (NF * N) / 64);
  uint32_t* bits = reinterpret_cast<uint32_t*>(npl + 16 * T);
  const int tid = int(threadIdx.x);
  const int W = a.W;
  const size_t HW = size_t(W) * N;
  const int img = a.img[blockIdx.y];
  const bool fresh = a.fresh != 0;
  // column pairs sharing 128-byte segments run on one XCD, one after the other
  const int nb = int(gridDim.x);
  int cp = int(blockIdx.x);
  if ((nb & 7) == 0) cp = (cp & 7) * (nb >> 3) + (cp >> 3);
  const int x0 = cp * NF;
  // the row bits of the pass's layers: bit li * tilesY + ty
  const int nl = a.nLayersInPass;
  const size_t nbits = size_t(nl) * a.tilesY;
  const size_t bit0 = size_t(a.passStart) * a.tilesY;
  const size_t w0 = bit0 >> 5;
  const int sh = int(bit0 & 31);
  for (size_t wd = size_t(tid); wd < (nbits + 31) / 32; wd += T) {
    uint32_t v = a.rowLive[w0 + wd] >> sh;
    if (sh != 0 && w0 + wd + 1 < a.rowLiveWords) v |= a.rowLive[w0 + wd + 1] << (32 - sh);
    bits[wd] = v;
  }
  // this thread's output pixels: n' and the accumulators
  float s[16], c[16], is[CPLX ? 16 : 1], ic[CPLX ? 16 : 1];
#pragma unroll
  for (int q = 0; q < PERL; ++q) {
    const int bt = tid + q * T;
    const int f = bt % NF, b = bt / NF;
    const int base = (b / NSL) * NSL * RL + b % NSL;
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int k = q * RL + r;
      const int y = base + r * NSL, x = x0 + f;
      int ySrc = (N - y) + N / 2;
      if (ySrc >= N) ySrc -= N;
      int xSrc = x + W / 2;
      if (xSrc >= W) xSrc -= W;
      const double m = pixelM(ySrc, N, a.dm, a.mShift);
      const double l = pixelL(xSrc, W, a.dl, a.lShift);
      double np = nPrime(l, m);
      if (REF) np = double(float(np));
      npl[k * T + tid] = np;
      const size_t o = size_t(img) * HW + size_t(ySrc) * W + xSrc;
      // (read also in a fresh pass, where the values are not used: no
      // branch around the loads)
      s[k] = fresh ? 0.0f : sum[o];
      c[k] = fresh ? 0.0f : comp[o];
      if (CPLX) {
        is[k] = fresh ? 0.0f : isum[o];
        ic[k] = fresh ? 0.0f : icomp[o];
      }
    }
  }
  __syncthreads();  // the row bits
  const int f0 = tid % NF, b0 = tid / NF;
  const float2* col = layers + size_t(img) * a.nLayersPass * HW + size_t(x0) * 4;
  float2 in[16], out[16];
  // cells of tile rows without samples were never written: their loads
  // take an offset out of the buffer's range and return zero (no branch
  // around the loads, and 32-bit offsets against a per-layer resource)
  const int colBytes = int((HW - size_t(x0) * 4) * sizeof(float2));
  auto load = [&](int li, float2* dst) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2*>(col + size_t(li) * HW), 0, colBytes, kRawBufferFlags);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int pos = b0 + r * B0;
      const int bit = li * a.tilesY + (pos >> a.tshift);
      const uint32_t off = ((bits[bit >> 5] >> (bit & 31)) & 1u)
                               ? uint32_t(blockedAddr<false>(pos, f0, W) * sizeof(float2))
                               : kOffsetOutOfRange;
      const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
      // (the elements by value: a bit cast of a vector element would read
      // element 0 for both with this compiler)
      const unsigned lo = v[0], hi = v[1];
      dst[r] = make_float2(__uint_as_float(lo), __uint_as_float(hi));
    }
  };
  load(0, in);
#pragma unroll 1
  for (int li = 0; li < nl; ++li) {
    // the next layer's inputs are loaded once the first stage has read this one's
    regFft<+1, N, NF>(lds, tw, tid, in, out, [&] {
      if (li + 1 < nl) load(li + 1, in);
    });
    const double w = layerToW(a.map, a.passStart + li);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float2 t = projectTerm<REF>(out[k], w, npl[k * T + tid]);
      kahanAdd(s[k], c[k], t.x);
      if (CPLX) kahanAdd(is[k], ic[k], t.y);
      // one cell's term at a time: interleaving the sixteen would hold
      // their temporaries at once
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  const bool zeroImag = !CPLX && fresh && isum != nullptr;
#pragma unroll
  for (int q = 0; q < PERL; ++q) {
    const int bt = tid + q * T;
    const int f = bt % NF, b = bt / NF;
    const int base = (b / NSL) * NSL * RL + b % NSL;
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int k = q * RL + r;
      const int y = base + r * NSL, x = x0 + f;
      int ySrc = (N - y) + N / 2;
      if (ySrc >= N) ySrc -= N;
      int xSrc = x + W / 2;
      if (xSrc >= W) xSrc -= W;
      const size_t o = size_t(img) * HW + size_t(ySrc) * W + xSrc;
      sum[o] = s[k];
      comp[o] = c[k];
      if (CPLX) {
        isum[o] = is[k];
        icomp[o] = ic[k];
      } else if (zeroImag) {
        // a fresh pass also zeroes the imaginary accumulators of real images
        isum[o] = 0.0f;
        icomp[o] = 0.0f;
      }
    }
  }
}

// Forward rows pass computed from the model for every layer of the batch
// (copyImageToLayerAndInverseCorrect, gridding/wstackinggridder.cpp:934-969,
// then the row transforms): a workgroup owns two rows of one image, reads
// their model values and computes their n' once, and then, layer after
// layer, forms each cell's w-phased input (modelTerm, as modelCell forms it),
// transforms the two rows and writes the tile columns the columns pass reads
// (colNeed: one bit per (layer, tile column); the columns pass skips the
// others). The cells' model values and n' stay in registers across the
// layers, so a layer costs the phase, the transform and the stores.
// grid (H / 2): row pair rp = rows 2 rp, 2 rp + 1, pairs of a four-row block
// run on one XCD one after the other (their stores fill the same lines).
template <int N, bool REF, bool CPLX>
__global__ void __launch_bounds__(kProjectColumns * N / 16)
fftRowsModelLayersKernel(float2* __restrict__ data, int H, int nbatch,
                         const float2* __restrict__ tw, ModelSource ms,
                         const uint32_t* __restrict__ colNeed, int tshift) {
  constexpr int NF = kProjectColumns;
  constexpr int T = NF * N / 16;
  constexpr int B0 = N / 16;
  constexpr int RL = LastStage<N>::R, NSL = LastStage<N>::NS, PERL = 16 / RL;
  __shared__ float2 lds[NF * N + (NF * N) / 64];
  const int tid = int(threadIdx.x);
  const int W = N;
  const size_t HW = size_t(W) * H;
  const int nb = int(gridDim.x);
  int rp = int(blockIdx.x);
  if ((nb & 7) == 0) rp = (rp & 7) * (nb >> 3) + (rp >> 3);
  const int y0 = rp * NF;
  const int tilesX = W >> tshift;
  // this thread's first-stage cells: x = b0 + r N / 16 of row y0 + f0
  const int f0 = tid % NF, b0 = tid / NF;
  float re[16], imv[CPLX ? 16 : 1];
  double np[16];
  {
    const int y = y0 + f0;
    int yDest = y + H / 2;
    if (yDest >= H) yDest -= H;
    const double m = pixelM(yDest, H, ms.dm, ms.mShift);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int x = b0 + r * B0;
      int xDest = (W - x) + W / 2;
      if (xDest >= W) xDest -= W;
      const double l = pixelL(xDest, W, ms.dl, ms.lShift);
      np[r] = nPrimeSeries(l, m);
      if (REF) np[r] = double(float(np[r]));
      const size_t src = size_t(yDest) * W + xDest;
      re[r] = ms.re[src];
      if (CPLX) imv[r] = -ms.im[src];
    }
  }
  // the rows' cells in the blocked layout: row y0 + f of a layer at
  // rowBase + layer offset + x * 4 + f
  const size_t rowBase = size_t(y0 >> 2) * 4 * W + size_t(y0 & 3);
  float2 in[16], out[16];
#pragma unroll 1
  for (int b = 0; b < nbatch; ++b) {
    const int layer = ms.layer0 + b;
    const double w = layerToW(ms.map, layer);
#pragma unroll
    for (int r = 0; r < 16; ++r)
      in[r] = modelTerm<REF>(re[r], CPLX ? imv[r] : 0.0f, CPLX, w, np[r]);
    regFft<-1, N, NF>(lds, tw, tid, in, out, [] {});
    float2* g = data + size_t(b) * HW + rowBase;
    const size_t cbit = size_t(layer) * tilesX;
#pragma unroll
    for (int q = 0; q < PERL; ++q) {
      const int bt = tid + q * T;
      const int f = bt % NF, bb = bt / NF;
      const int base = (bb / NSL) * NSL * RL + bb % NSL;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int x = base + r * NSL;
        const size_t bit = cbit + size_t(x >> tshift);
        if ((colNeed[bit >> 5] >> (bit & 31)) & 1u)
          streamStore(&g[size_t(x) * 4 + f], out[q * RL + r]);
      }
    }
  }
}

// rows of length N, NFFT rows per block (the last block may hold fewer)
template <int SIGN, int N, int NT, int NFFT>
__global__ void __launch_bounds__(NT)
fftRowsPow2Kernel(float2* __restrict__ data, int H, const float2* __restrict__ tw) {
  __shared__ float2 lds[NFFT * N];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.x * NFFT;
  const int nrows = min(NFFT, H - row0);
  float2* base = data + size_t(blockIdx.y) * size_t(N) * size_t(H) + size_t(row0) * N;
  const int total = nrows * N;
  for (int i = tid; i < total; i += NT) lds[i] = base[i];
  __syncthreads();
  pow2Fft<SIGN, N, NT, NFFT>(lds, N, nrows, tw, tid);
  for (int i = tid; i < total; i += NT) base[i] = lds[i];
}

// columns of length N (= H), NFFT columns per block, LDS [NFFT][N + 1]
template <int SIGN, int N, int NT, int NFFT>
__global__ void __launch_bounds__(NT)
fftColsPow2Kernel(float2* __restrict__ data, int W, int nColBlocks,
                  const float2* __restrict__ tw) {
  __shared__ float2 lds[NFFT * (N + 1)];
  const int tid = threadIdx.x;
  int blk = blockIdx.x;
  if ((nColBlocks & 7) == 0) blk = (blk & 7) * (nColBlocks >> 3) + (blk >> 3);
  const int c0 = blk * NFFT;
  const int ncols = min(NFFT, W - c0);
  float2* base = data + size_t(blockIdx.y) * size_t(W) * size_t(N) + c0;
  const int total = ncols * N;
  for (int i = tid; i < total; i += NT) {
    const int y = i / ncols, c = i - y * ncols;
    lds[c * (N + 1) + y] = base[size_t(y) * W + c];
  }
  __syncthreads();
  pow2Fft<SIGN, N, NT, NFFT>(lds, N + 1, ncols, tw, tid);
  for (int i = tid; i < total; i += NT) {
    const int y = i / ncols, c = i - y * ncols;
    base[size_t(y) * W + c] = lds[c * (N + 1) + y];
  }
}

// rows: grid (ceil(H / rowsPerBlock), nbatch); each block transforms
// rowsPerBlock consecutive rows of one layer.
template <int SIGN, int NT>
__global__ void __launch_bounds__(NT)
fftRowsKernel(float2* __restrict__ data, int W, int H, int rowsPerBlock,
              unsigned long long radices, const float2* __restrict__ tw) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int row0 = blockIdx.x * rowsPerBlock;
  const int nrows = min(rowsPerBlock, H - row0);
  float2* base = data + size_t(blockIdx.y) * size_t(W) * size_t(H) +
                 size_t(row0) * size_t(W);
  const int pitch = ldsPitch(W);
  for (int r = 0; r < nrows; ++r)
    for (int x = tid; x < W; x += nthr) lds[r * pitch + lp(x)] = base[size_t(r) * W + x];
  __syncthreads();
  fftLds<SIGN>(lds, pitch, nrows, W, radices, tw, tid, nthr);
  for (int r = 0; r < nrows; ++r)
    for (int x = tid; x < W; x += nthr) base[size_t(r) * W + x] = lds[r * pitch + lp(x)];
}

// columns: grid (nColBlocks, nbatch); block handles columns
// [c0, c0 + bc) of one layer; LDS holds [bc][ldsPitch(H)].
template <int SIGN, int NT>
__global__ void __launch_bounds__(NT)
fftColsKernel(float2* __restrict__ data, int W, int H, int bc, int nColBlocks,
              unsigned long long radices, const float2* __restrict__ tw) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  // XCD-aware remap: hardware dispatches block b to XCD b % 8; give each XCD
  // a contiguous range of column blocks so neighbours share its L2.
  int blk = blockIdx.x;
  if ((nColBlocks & 7) == 0) blk = (blk & 7) * (nColBlocks >> 3) + (blk >> 3);
  const int c0 = blk * bc;
  const int ncols = min(bc, W - c0);
  const int pitch = ldsPitch(H);
  float2* base = data + size_t(blockIdx.y) * size_t(W) * size_t(H) + c0;
  const int total = ncols * H;
  for (int i = tid; i < total; i += nthr) {
    const int y = i / ncols, c = i - y * ncols;
    lds[c * pitch + lp(y)] = base[size_t(y) * W + c];
  }
  __syncthreads();
  fftLds<SIGN>(lds, pitch, ncols, H, radices, tw, tid, nthr);
  for (int i = tid; i < total; i += nthr) {
    const int y = i / ncols, c = i - y * ncols;
    base[size_t(y) * W + c] = lds[c * pitch + lp(y)];
  }
}

// Blocked layout at sizes without a compile-time kernel (any 7-smooth size
// with W and H multiples of 4): one workgroup per 4 rows (row pass) or 4
// columns (column pass), read and written as whole 128-byte segments of the
// blocked layout (two rows or columns, 64-byte segments, where four do not
// fit the LDS), by the mixed-radix stages: the first reads memory, the middle
// ones exchange through the lines interleaved in LDS, the last writes memory.
// The sparse modes mirror the compile-time kernels:
//   rows MODE 1 (backward): only tiles with samples are read (bits: the
//     occupancy of (layer, tile), sp.occ); a row block without any is
//     skipped (not written); MODE 2 (forward): the input is computed from the
//     model image (ms);
//   columns MODE 1 (backward): tile rows without samples (sp.rowLive) read as
//     zero; MODE 2 (forward): only the tiles degridding reads (sp.occ = the
//     need bits) are written, and a column block without any is skipped.
struct GenericSparseArgs {
  const uint32_t* bits;
  int layer0, tilesX, tilesY, tshift;
};

__device__ __forceinline__ bool bitAt(const uint32_t* bits, size_t b) {
  return (bits[b >> 5] >> (b & 31)) & 1u;
}

// LDS position of value i of the interleaved lines of a blocked runtime-radix
// workgroup (element e of line f is value e * L + f): one pad value after
// every 16, so that the first stage's writes (stride R * L values across the
// lanes) fall in distinct banks.
__device__ __forceinline__ int ipad(int i) { return i + (i >> 4); }

// LDS bytes of L interleaved lines of length n.
__host__ __device__ constexpr size_t blockedGenericLds(int L, int n) {
  return (size_t(L) * n + (size_t(L) * n >> 4) + 1) * sizeof(float2);
}

// One Stockham stage of radix R over the L interleaved lines of length n
// (butterfly t: line f = t % L, index b = t / L, so the lanes of a wave read
// and write the L lines of consecutive positions together). Butterfly b reads
// x[b + r*B] (B = n/R), applies twiddle w^(jm*r) with jm = b % ns, and writes
// y[(b/ns)*ns*R + jm + r*ns]. The first stage (ns == 1) reads its inputs with
// load(slot, pos, f) instead of from LDS (slot = q * R + r, this thread's
// q-th butterfly's r-th input, below kMaxValues) and needs no barrier before
// its writes; the last one writes its outputs with store(slot, pos, f, v)
// instead of to LDS.
template <int SIGN, int R, int L, bool FIRST, bool LAST, class Load, class Store>
__device__ __forceinline__ void blockedRtStage(float2* lds, int n, int ns,
                                               const float2* __restrict__ tw, int tid, int nthr,
                                               const Load& load, const Store& store) {
  constexpr int kBatches = kMaxValues / R;
  constexpr int LS = L == 4 ? 2 : L == 2 ? 1 : 0;
  static_assert(L == 1 || L == 2 || L == 4, "lines");
  // (an opaque tid keeps the compiler from hoisting the batch indices of
  // every radix out of the stage loop, which keeps them all live)
  asm volatile("" : "+v"(tid));
  const int B = n / R;
  const int total = B * L;
  const int twStep = n / (ns * R);
  const UniformDiv byNs = uniformDiv(ns);
  float2 v[kBatches][R];
#pragma unroll
  for (int q = 0; q < kBatches; ++q) {
    const int t = tid + q * nthr;
    if (t < total) {
      const int f = t & (L - 1), b = t >> LS;
      const int jm = FIRST ? 0 : b - divBy(b, byNs) * ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int pos = b + r * B;
        float2 x;
        if (FIRST) {
          x = load(q * R + r, pos, f);
        } else {
          x = lds[ipad(pos * L + f)];
          if (r != 0 && jm != 0) x = cmul(x, tw[jm * r * twStep]);
        }
        v[q][r] = x;
      }
    }
  }
  if (!FIRST) __syncthreads();  // every read of this stage precedes its writes
#pragma unroll
  for (int q = 0; q < kBatches; ++q) {
    const int t = tid + q * nthr;
    if (t < total) {
      const int f = t & (L - 1), b = t >> LS;
      Butterfly<SIGN, R>::run(v[q]);
      const int bq = FIRST ? b : divBy(b, byNs);
      const int idxD = bq * ns * R + (b - bq * ns);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int pos = idxD + r * ns;
        if (LAST)
          store(q * R + r, pos, f, v[q][r]);
        else
          lds[ipad(pos * L + f)] = v[q][r];
      }
    }
  }
  if (!LAST) __syncthreads();
}

template <int SIGN, int L, bool FIRST, bool LAST, class Load, class Store>
__device__ __forceinline__ void blockedRtStageR(int R, float2* lds, int n, int ns,
                                                const float2* __restrict__ tw, int tid, int nthr,
                                                const Load& load, const Store& store) {
  switch (R) {
#define HWS_RT_STAGE(RR)                                                                    \
  case RR:                                                                                  \
    blockedRtStage<SIGN, RR, L, FIRST, LAST>(lds, n, ns, tw, tid, nthr, load, store);       \
    break;
    HWS_RT_STAGE(16)
    HWS_RT_STAGE(8)
    HWS_RT_STAGE(4)
    HWS_RT_STAGE(2)
    HWS_RT_STAGE(3)
    HWS_RT_STAGE(5)
    HWS_RT_STAGE(7)
#undef HWS_RT_STAGE
    default: break;
  }
}

// The whole transform of the L lines: the first stage from load, the middle
// stages through LDS, the last stage to store (radices packed 4 bits per
// stage, 16 packed as 1).
template <int SIGN, int L, class Load, class Store>
__device__ __forceinline__ void blockedRtFft(float2* lds, int n, unsigned long long radices,
                                             const float2* __restrict__ tw, int tid, int nthr,
                                             const Load& load, const Store& store) {
  auto radixOf = [](unsigned long long rad) {
    const int code = int(rad & 15ull);
    return code == 1 ? 16 : code;
  };
  int R = radixOf(radices);
  radices >>= 4;
  auto toLds = [](int, int, int, float2) {};
  auto fromLds = [](int, int, int) { return make_float2(0.0f, 0.0f); };
  if (radices == 0ull) {  // a single stage: memory to memory
    blockedRtStageR<SIGN, L, true, true>(R, lds, n, 1, tw, tid, nthr, load, store);
    return;
  }
  blockedRtStageR<SIGN, L, true, false>(R, lds, n, 1, tw, tid, nthr, load, toLds);
  int ns = R;
  for (;;) {
    R = radixOf(radices);
    radices >>= 4;
    if (radices == 0ull) break;
    blockedRtStageR<SIGN, L, false, false>(R, lds, n, ns, tw, tid, nthr, fromLds, toLds);
    ns *= R;
  }
  blockedRtStageR<SIGN, L, false, true>(R, lds, n, ns, tw, tid, nthr, fromLds, store);
}

// Fewer than four lines per workgroup: the workgroups of a four-line block
// read and write parts of the same 32-byte (rows) or 128-byte (columns)
// segments. The hardware sends workgroup b to XCD b % 8, so within every
// 8 * P workgroups (P = 4 / L parts per block) b, b + 8, ... are given the
// parts of one block: they run on the same XCD at about the same time and
// share its L2.
template <int L>
__device__ __forceinline__ int sharedSegmentBlock(int b, int nBlocks) {
  constexpr int P = 4 / L, G = 8 * P;
  if (P == 1 || nBlocks % G != 0) return b;
  return (b / G) * G + (b & 7) * P + ((b >> 3) & (P - 1));
}

template <int SIGN, int MODE, bool REF, int L>
__global__ void __launch_bounds__(1024)
fftRowsBlockedGenericKernel(float2* __restrict__ data, int W, int H,
                            unsigned long long radices, const float2* __restrict__ tw,
                            GenericSparseArgs sp, ModelSource ms) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int rblk = sharedSegmentBlock<L>(int(blockIdx.x), int(gridDim.x));
  const int r0 = rblk * L;                // rows of this workgroup: 4, 2 or 1 of a row block
  const int f0 = r0 & 3;                  // their position in the four-row block
  const int layer = sp.layer0 + int(blockIdx.y);
  const int ty = r0 >> sp.tshift;
  const size_t rowBase = (size_t(layer) * sp.tilesY + ty) * sp.tilesX;
  if (MODE == 1) {
    int any = 0;
    for (int t = tid; t < sp.tilesX; t += nthr) any |= bitAt(sp.bits, rowBase + t);
    if (!__syncthreads_or(any)) return;
  }
  float2* base = data + size_t(blockIdx.y) * size_t(W) * size_t(H) + size_t(r0 >> 2) * 4 * W + f0;
  const double w = MODE == 2 ? layerToW(ms.map, ms.layer0 + int(blockIdx.y)) : 0.0;
  auto load = [&](int, int x, int f) -> float2 {
    if (MODE == 2) return modelCell<REF>(ms, w, x, r0 + f, W, H);
    if (MODE == 1 && !bitAt(sp.bits, rowBase + (x >> sp.tshift))) return make_float2(0.0f, 0.0f);
    return base[size_t(x) * 4 + f];
  };
  auto store = [&](int, int x, int f, float2 v) { base[size_t(x) * 4 + f] = v; };
  blockedRtFft<SIGN, L>(lds, W, radices, tw, tid, nthr, load, store);
}

template <int SIGN, int MODE, int L>
__global__ void __launch_bounds__(1024)
fftColsBlockedGenericKernel(float2* __restrict__ data, int W, int H, int nColBlocks,
                            unsigned long long radices, const float2* __restrict__ tw,
                            GenericSparseArgs sp) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  int blk = blockIdx.x;
  // (the sparse forward mode keeps the natural order: its live column blocks
  // lie near the uv origin, at both ends of the row)
  if (MODE != 2 && (nColBlocks & 7) == 0)
    blk = (blk & 7) * (nColBlocks >> 3) + (blk >> 3);
  else if (MODE == 2)
    blk = sharedSegmentBlock<L>(blk, nColBlocks);
  const int layer = sp.layer0 + int(blockIdx.y);
  const int c0 = blk * L;  // columns of this workgroup: 4, 2 or 1
  const int tx = c0 >> sp.tshift;
  // MODE 1: bit (layer, tile row); MODE 2: bit (layer, tile row, this tile column)
  auto live = [&](int y) -> bool {
    const int ty = y >> sp.tshift;
    if (MODE == 1) return bitAt(sp.bits, size_t(layer) * sp.tilesY + ty);
    return bitAt(sp.bits, (size_t(layer) * sp.tilesY + ty) * sp.tilesX + tx);
  };
  if (MODE == 2) {
    int any = 0;
    for (int t = tid; t < sp.tilesY; t += nthr) any |= live(t << sp.tshift);
    if (!__syncthreads_or(any)) return;
  }
  float2* base = data + size_t(blockIdx.y) * size_t(W) * size_t(H) + size_t(c0) * 4;
  // element (y, column c0 + f): row block y >> 2, row y & 3 in it
  auto addr = [&](int y, int f) { return size_t(y >> 2) * W * 4 + size_t(f) * 4 + (y & 3); };
  auto load = [&](int, int y, int f) -> float2 {
    if (MODE == 1 && !live(y)) return make_float2(0.0f, 0.0f);
    return base[addr(y, f)];
  };
  auto store = [&](int, int y, int f, float2 v) {
    if (MODE == 2 && !live(y)) return;
    base[addr(y, f)] = v;
  };
  blockedRtFft<SIGN, L>(lds, H, radices, tw, tid, nthr, load, store);
}

// The first-stage cells of this thread in a blocked runtime-radix transform
// whose first radix is R: visit(slot, pos, f) for each (the slots of
// blockedRtStage's first stage).
template <int R, int L, class Visit>
__device__ __forceinline__ void firstStageCells(int n, int tid, int nthr, const Visit& visit) {
  constexpr int kBatches = kMaxValues / R;
  constexpr int LS = L == 4 ? 2 : L == 2 ? 1 : 0;
  asm volatile("" : "+v"(tid));
  const int B = n / R;
  const int total = B * L;
#pragma unroll
  for (int q = 0; q < kBatches; ++q) {
    const int t = tid + q * nthr;
    if (t < total) {
      const int f = t & (L - 1), b = t >> LS;
#pragma unroll
      for (int r = 0; r < R; ++r) visit(q * R + r, b + r * B, f);
    }
  }
}

// Forward rows pass computed from the model for every layer of the batch at
// the sizes the runtime-radix kernels serve (fftRowsModelLayersKernel at the
// power-of-two sizes): a workgroup owns L rows of one image, reads their
// model values and computes their n' once into registers (the cells of its
// first stage), and then, layer after layer, forms each cell's w-phased
// input (modelTerm, as modelCell forms it), transforms the rows and writes
// the tile columns the columns pass reads (colNeed: one bit per (layer, tile
// column)). grid (H / L).
template <bool REF, bool CPLX, int L>
__global__ void __launch_bounds__(1024)
fftRowsModelLayersGenericKernel(float2* __restrict__ data, int W, int H, int nbatch,
                                unsigned long long radices, const float2* __restrict__ tw,
                                ModelSource ms, const uint32_t* __restrict__ colNeed,
                                int tshift) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int rblk = sharedSegmentBlock<L>(int(blockIdx.x), int(gridDim.x));
  const int r0 = rblk * L;
  const size_t HW = size_t(W) * H;
  float2* rowBase = data + size_t(r0 >> 2) * 4 * W + (r0 & 3);
  const int tilesX = W >> tshift;
  float re[kMaxValues], imv[CPLX ? kMaxValues : 1];
  double np[kMaxValues];
  auto cell = [&](int slot, int x, int f) {
    const int y = r0 + f;
    int yDest = y + H / 2;
    if (yDest >= H) yDest -= H;
    int xDest = (W - x) + W / 2;
    if (xDest >= W) xDest -= W;
    const double m = pixelM(yDest, H, ms.dm, ms.mShift);
    const double l = pixelL(xDest, W, ms.dl, ms.lShift);
    double v = nPrimeSeries(l, m);
    if (REF) v = double(float(v));
    np[slot] = v;
    const size_t src = size_t(yDest) * W + xDest;
    re[slot] = ms.re[src];
    if (CPLX) imv[slot] = -ms.im[src];
  };
  switch (radices & 15ull) {
    case 1: firstStageCells<16, L>(W, tid, nthr, cell); break;
    case 8: firstStageCells<8, L>(W, tid, nthr, cell); break;
    case 4: firstStageCells<4, L>(W, tid, nthr, cell); break;
    case 2: firstStageCells<2, L>(W, tid, nthr, cell); break;
    case 3: firstStageCells<3, L>(W, tid, nthr, cell); break;
    case 5: firstStageCells<5, L>(W, tid, nthr, cell); break;
    case 7: firstStageCells<7, L>(W, tid, nthr, cell); break;
    default: break;
  }
#pragma unroll 1
  for (int b = 0; b < nbatch; ++b) {
    const int layer = ms.layer0 + b;
    const double w = layerToW(ms.map, layer);
    float2* base = rowBase + size_t(b) * HW;
    const size_t cbit = size_t(layer) * tilesX;
    auto load = [&](int slot, int, int) -> float2 {
      return modelTerm<REF>(re[slot], CPLX ? imv[slot] : 0.0f, CPLX, w, np[slot]);
    };
    auto store = [&](int, int x, int f, float2 v) {
      const size_t bit = cbit + size_t(x >> tshift);
      if ((colNeed[bit >> 5] >> (bit & 31)) & 1u) streamStore(&base[size_t(x) * 4 + f], v);
    };
    blockedRtFft<-1, L>(lds, W, radices, tw, tid, nthr, load, store);
    __syncthreads();  // the last stage's LDS reads precede the next layer's writes
  }
}

template <int SIGN, int MODE, bool REF>
void launchRowsGeneric(int threads, int lines, dim3 grid, size_t lds, hipStream_t st,
                       float2* data, int W, int H, unsigned long long rad, const float2* tw,
                       GenericSparseArgs sp, ModelSource ms) {
  if (lines == 4)
    hipLaunchKernelGGL((fftRowsBlockedGenericKernel<SIGN, MODE, REF, 4>), grid, dim3(threads),
                       lds, st, data, W, H, rad, tw, sp, ms);
  else if (lines == 2)
    hipLaunchKernelGGL((fftRowsBlockedGenericKernel<SIGN, MODE, REF, 2>), grid, dim3(threads),
                       lds, st, data, W, H, rad, tw, sp, ms);
  else
    hipLaunchKernelGGL((fftRowsBlockedGenericKernel<SIGN, MODE, REF, 1>), grid, dim3(threads),
                       lds, st, data, W, H, rad, tw, sp, ms);
}

// Backward columns pass fused with the projection at the runtime-radix
// sizes (fftColsProjectKernel at the power-of-two sizes): a workgroup owns
// one column of one image and transforms it layer after layer, in the pass's
// layer order; each transformed cell is rotated by its layer's w term and
// added (Kahan) to its pixel's sums, which stay in registers with the
// pixel's n' (the cells of the last stage's slots). Cells of tile rows
// without samples read as zero. Every term is computed and added as
// projectKernel computes and adds it. grid (W / 1, images of the kind);
// dynamic LDS: the transform buffer, then the row bits of the pass's layers.
template <bool REF, bool CPLX>
__global__ void __launch_bounds__(512)
fftColsProjectGenericKernel(ColsProjectArgs a, const float2* __restrict__ layers,
                            unsigned long long radices, const float2* __restrict__ tw,
                            float* __restrict__ sum, float* __restrict__ comp,
                            float* __restrict__ isum, float* __restrict__ icomp) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int W = a.W, H = a.H;
  const size_t HW = size_t(W) * H;
  const int img = a.img[blockIdx.y];
  const bool fresh = a.fresh != 0;
  // the four columns of a 128-byte segment run on one XCD, one after the other
  const int nb = int(gridDim.x);
  int x0 = int(blockIdx.x);
  if ((nb & 7) == 0) x0 = (x0 & 7) * (nb >> 3) + (x0 >> 3);
  uint32_t* bits =
      reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(lds) + blockedGenericLds(1, H));
  const int nl = a.nLayersInPass;
  {
    const size_t nbits = size_t(nl) * a.tilesY;
    const size_t bit0 = size_t(a.passStart) * a.tilesY;
    const size_t w0 = bit0 >> 5;
    const int sh = int(bit0 & 31);
    for (size_t wd = size_t(tid); wd < (nbits + 31) / 32; wd += size_t(nthr)) {
      uint32_t v = a.rowLive[w0 + wd] >> sh;
      if (sh != 0 && w0 + wd + 1 < a.rowLiveWords) v |= a.rowLive[w0 + wd + 1] << (32 - sh);
      bits[wd] = v;
    }
  }
  // this thread's output pixels (the last stage's slots): n' and the sums
  float s[kMaxValues], c[kMaxValues], is[CPLX ? kMaxValues : 1], ic[CPLX ? kMaxValues : 1];
  double np[kMaxValues];
  auto pixel = [&](int y) {
    int ySrc = (H - y) + H / 2;
    if (ySrc >= H) ySrc -= H;
    int xSrc = x0 + W / 2;
    if (xSrc >= W) xSrc -= W;
    return size_t(img) * HW + size_t(ySrc) * W + xSrc;
  };
  auto init = [&](int slot, int y, int) {
    int ySrc = (H - y) + H / 2;
    if (ySrc >= H) ySrc -= H;
    int xSrc = x0 + W / 2;
    if (xSrc >= W) xSrc -= W;
    const double m = pixelM(ySrc, H, a.dm, a.mShift);
    const double l = pixelL(xSrc, W, a.dl, a.lShift);
    double v = nPrime(l, m);
    if (REF) v = double(float(v));
    np[slot] = v;
    const size_t o = pixel(y);
    s[slot] = fresh ? 0.0f : sum[o];
    c[slot] = fresh ? 0.0f : comp[o];
    if (CPLX) {
      is[slot] = fresh ? 0.0f : isum[o];
      ic[slot] = fresh ? 0.0f : icomp[o];
    }
  };
  auto finish = [&](int slot, int y, int) {
    const size_t o = pixel(y);
    sum[o] = s[slot];
    comp[o] = c[slot];
    if (CPLX) {
      isum[o] = is[slot];
      icomp[o] = ic[slot];
    } else if (fresh && isum != nullptr) {
      // a fresh pass also zeroes the imaginary accumulators of real images
      isum[o] = 0.0f;
      icomp[o] = 0.0f;
    }
  };
  // the last stage's radix: the highest nonzero 4-bit code
  int lastCode = 0;
  for (unsigned long long r = radices; r != 0ull; r >>= 4) lastCode = int(r & 15ull);
  auto overLastSlots = [&](const auto& visit) {
    switch (lastCode) {
      case 1: firstStageCells<16, 1>(H, tid, nthr, visit); break;
      case 8: firstStageCells<8, 1>(H, tid, nthr, visit); break;
      case 4: firstStageCells<4, 1>(H, tid, nthr, visit); break;
      case 2: firstStageCells<2, 1>(H, tid, nthr, visit); break;
      case 3: firstStageCells<3, 1>(H, tid, nthr, visit); break;
      case 5: firstStageCells<5, 1>(H, tid, nthr, visit); break;
      case 7: firstStageCells<7, 1>(H, tid, nthr, visit); break;
      default: break;
    }
  };
  overLastSlots(init);
  __syncthreads();  // the row bits
  const float2* col = layers + size_t(img) * a.nLayersPass * HW + size_t(x0) * 4;
#pragma unroll 1
  for (int li = 0; li < nl; ++li) {
    const float2* lay = col + size_t(li) * HW;
    const double w = layerToW(a.map, a.passStart + li);
    auto load = [&](int, int y, int) -> float2 {
      const int bit = li * a.tilesY + (y >> a.tshift);
      if (!((bits[bit >> 5] >> (bit & 31)) & 1u)) return make_float2(0.0f, 0.0f);
      return lay[size_t(y >> 2) * W * 4 + (y & 3)];
    };
    auto project = [&](int slot, int, int, float2 v) {
      const float2 t = projectTerm<REF>(v, w, np[slot]);
      kahanAdd(s[slot], c[slot], t.x);
      if (CPLX) kahanAdd(is[slot], ic[slot], t.y);
    };
    blockedRtFft<+1, 1>(lds, H, radices, tw, tid, nthr, load, project);
    __syncthreads();  // the last stage's LDS reads precede the next layer's writes
  }
  overLastSlots(finish);
}

void launchModelLayersGeneric(int threads, int lines, dim3 grid, size_t lds, hipStream_t st,
                              float2* data, int W, int H, int nbatch, unsigned long long rad,
                              const float2* tw, const ModelSource& ms, const uint32_t* colNeed,
                              int tshift) {
  const bool cplx = ms.cplx != 0;
#define HWS_MODEL_GENERIC(C, LL)                                                                \
  hipLaunchKernelGGL((fftRowsModelLayersGenericKernel<false, C, LL>), grid, dim3(threads), lds,  \
                     st, data, W, H, nbatch, rad, tw, ms, colNeed, tshift)
  if (lines == 4) {
    if (cplx) HWS_MODEL_GENERIC(true, 4);
    else HWS_MODEL_GENERIC(false, 4);
  } else if (lines == 2) {
    if (cplx) HWS_MODEL_GENERIC(true, 2);
    else HWS_MODEL_GENERIC(false, 2);
  } else {
    if (cplx) HWS_MODEL_GENERIC(true, 1);
    else HWS_MODEL_GENERIC(false, 1);
  }
#undef HWS_MODEL_GENERIC
}

template <int SIGN, int MODE>
void launchColsGeneric(int threads, int lines, dim3 grid, size_t lds, hipStream_t st,
                       float2* data, int W, int H, int nColBlocks, unsigned long long rad,
                       const float2* tw, GenericSparseArgs sp) {
  if (lines == 4)
    hipLaunchKernelGGL((fftColsBlockedGenericKernel<SIGN, MODE, 4>), grid, dim3(threads), lds,
                       st, data, W, H, nColBlocks, rad, tw, sp);
  else if (lines == 2)
    hipLaunchKernelGGL((fftColsBlockedGenericKernel<SIGN, MODE, 2>), grid, dim3(threads), lds,
                       st, data, W, H, nColBlocks, rad, tw, sp);
  else
    hipLaunchKernelGGL((fftColsBlockedGenericKernel<SIGN, MODE, 1>), grid, dim3(threads), lds,
                       st, data, W, H, nColBlocks, rad, tw, sp);
}

// The plain (row-major) runtime-radix kernels are compiled for three launch
// bounds; a launch takes the smallest that holds its threads.
template <int SIGN>
void launchRowsPlain(int threads, dim3 grid, size_t lds, hipStream_t st, float2* data, int W,
                     int H, int rows, unsigned long long rad, const float2* tw) {
  if (threads <= 256)
    hipLaunchKernelGGL((fftRowsKernel<SIGN, 256>), grid, dim3(threads), lds, st, data, W, H, rows,
                       rad, tw);
  else if (threads <= 512)
    hipLaunchKernelGGL((fftRowsKernel<SIGN, 512>), grid, dim3(threads), lds, st, data, W, H, rows,
                       rad, tw);
  else
    hipLaunchKernelGGL((fftRowsKernel<SIGN, 1024>), grid, dim3(threads), lds, st, data, W, H,
                       rows, rad, tw);
}

template <int SIGN>
void launchColsPlain(int threads, dim3 grid, size_t lds, hipStream_t st, float2* data, int W,
                     int H, int bc, int nColBlocks, unsigned long long rad, const float2* tw) {
  if (threads <= 256)
    hipLaunchKernelGGL((fftColsKernel<SIGN, 256>), grid, dim3(threads), lds, st, data, W, H, bc,
                       nColBlocks, rad, tw);
  else if (threads <= 512)
    hipLaunchKernelGGL((fftColsKernel<SIGN, 512>), grid, dim3(threads), lds, st, data, W, H, bc,
                       nColBlocks, rad, tw);
  else
    hipLaunchKernelGGL((fftColsKernel<SIGN, 1024>), grid, dim3(threads), lds, st, data, W, H, bc,
                       nColBlocks, rad, tw);
}

unsigned long long packRadices(const FftPlan& plan) {
  unsigned long long r = 0;
  for (int i = plan.nstages - 1; i >= 0; --i)
    r = (r << 4) | unsigned(plan.radix[i] == 16 ? 1 : plan.radix[i]);
  return r;
}

// Threads such that every stage's butterflies fit in kMaxValues / R batches
// per thread; 0 if more than 1024 would be needed.
int threadsFor(const FftPlan& plan, int nfft) {
  const int elements = nfft * plan.n;
  for (int t = 64; t <= 1024; t *= 2) {
    bool ok = true;
    for (int s = 0; s < plan.nstages; ++s) {
      const int R = plan.radix[s];
      if (t * (kMaxValues / R) < elements / R) ok = false;
    }
    if (ok) return std::max(t, 64);
  }
  return 0;
}

}  // namespace

hws_status makeFftPlan(hws_gridder* g, int n, FftPlan& plan) {
  plan.n = n;
  plan.nstages = 0;
  int m = n;
  // radix 16 first, then 8 / 4 / 2, then the odd factors (fewer stages, each
  // a round trip through LDS)
  while (m % 16 == 0 && plan.nstages < kMaxFftStages) { plan.radix[plan.nstages++] = 16; m /= 16; }
  while (m % 8 == 0 && plan.nstages < kMaxFftStages) { plan.radix[plan.nstages++] = 8; m /= 8; }
  while (m % 4 == 0 && plan.nstages < kMaxFftStages) { plan.radix[plan.nstages++] = 4; m /= 4; }
  while (m % 2 == 0 && plan.nstages < kMaxFftStages) { plan.radix[plan.nstages++] = 2; m /= 2; }
  for (int r : {3, 5, 7})
    while (m % r == 0 && plan.nstages < kMaxFftStages) { plan.radix[plan.nstages++] = r; m /= r; }
  if (m != 1)
    return setLastError(g, HWS_ERR_UNSUPPORTED,
                        "image dimension " + std::to_string(n) +
                            " is not a product of 2, 3, 5 and 7");
  if (n > 16384 || plan.nstages > 16)
    return setLastError(g, HWS_ERR_UNSUPPORTED,
                        "image dimension above 16384 is not supported");
  std::vector<float2> tb(n), tf(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; ++k) {
    const long double a = 2.0L * pi * (long double)k / (long double)n;
    const float c = (float)cosl(a), s = (float)sinl(a);
    tb[k] = make_float2(c, s);
    tf[k] = make_float2(c, -s);
  }
  hws_status st;
  if ((st = checkHip(g, hipMalloc(&plan.twBackward, n * sizeof(float2)), "hipMalloc twiddles")) != HWS_OK) return st;
  if ((st = checkHip(g, hipMalloc(&plan.twForward, n * sizeof(float2)), "hipMalloc twiddles")) != HWS_OK) return st;
  if ((st = checkHip(g, hipMemcpy(plan.twBackward, tb.data(), n * sizeof(float2), hipMemcpyHostToDevice), "twiddle upload")) != HWS_OK) return st;
  return checkHip(g, hipMemcpy(plan.twForward, tf.data(), n * sizeof(float2), hipMemcpyHostToDevice), "twiddle upload");
}

void freeFftPlan(FftPlan& plan) {
  if (plan.twBackward) (void)hipFree(plan.twBackward);
  if (plan.twForward) (void)hipFree(plan.twForward);
  plan.twBackward = plan.twForward = nullptr;
}

namespace {

// Power-of-two launch shapes: rows take max(1, 4096/N) rows per block,
// columns max(1, 8192/N) columns (at most 16); 16 values per thread.
template <int SIGN, int N>
void launchPow2(float2* data, int W, int H, int nbatch, const float2* twRow,
                const float2* twCol, bool rows, hipStream_t stream) {
  if (rows) {
    constexpr int NFFT = (4096 / N) > 0 ? (4096 / N) : 1;
    constexpr int E = NFFT * N;
    constexpr int NT = (E / 16) < 64 ? 64 : (E / 16) > 1024 ? 1024 : (E / 16);
    const int nfft = std::min(NFFT, H);
    (void)nfft;
    dim3 grid((H + NFFT - 1) / NFFT, nbatch);
    hipLaunchKernelGGL((fftRowsPow2Kernel<SIGN, N, NT, NFFT>), grid, dim3(NT), 0, stream, data, H, twRow);
  } else {
    constexpr int NFFT0 = (8192 / N) > 0 ? (8192 / N) : 1;
    constexpr int NFFT = NFFT0 > 16 ? 16 : NFFT0;
    constexpr int E = NFFT * N;
    constexpr int NT = (E / 16) < 64 ? 64 : (E / 16) > 1024 ? 1024 : (E / 16);
    const int nColBlocks = (W + NFFT - 1) / NFFT;
    dim3 grid(nColBlocks, nbatch);
    hipLaunchKernelGGL((fftColsPow2Kernel<SIGN, N, NT, NFFT>), grid, dim3(NT), 0, stream, data, W, nColBlocks, twCol);
  }
}

template <int SIGN>
bool dispatchPow2(int n, float2* data, int W, int H, int nbatch, const float2* twRow,
                  const float2* twCol, bool rows, hipStream_t stream) {
  switch (n) {
    case 64: launchPow2<SIGN, 64>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 128: launchPow2<SIGN, 128>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 256: launchPow2<SIGN, 256>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 512: launchPow2<SIGN, 512>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 1024: launchPow2<SIGN, 1024>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 2048: launchPow2<SIGN, 2048>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 4096: launchPow2<SIGN, 4096>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    case 8192: launchPow2<SIGN, 8192>(data, W, H, nbatch, twRow, twCol, rows, stream); return true;
    default: return false;
  }
}

}  // namespace

namespace {
template <int SIGN>
bool launchBlocked(int n, bool rows, float2* data, int W, int H, int nbatch,
                   const float2* tw, hipStream_t stream) {
#define HWS_BLOCKED_CASE(NN)                                                            \
  case NN:                                                                              \
    if (rows)                                                                           \
      hipLaunchKernelGGL((fftRowsBlockedKernel<SIGN, NN>), dim3(H / 4, nbatch),         \
                         dim3(NN / 4), 0, stream, data, H, tw);                         \
    else                                                                                \
      hipLaunchKernelGGL((fftColsBlockedKernel<SIGN, NN>), dim3(W / 4, nbatch),         \
                         dim3(NN / 4), 0, stream, data, W, tw);                         \
    return true;
  switch (n) {
    HWS_BLOCKED_CASE(256)
    HWS_BLOCKED_CASE(512)
    HWS_BLOCKED_CASE(1024)
    HWS_BLOCKED_CASE(2048)
    HWS_BLOCKED_CASE(4096)
    default: return false;
  }
#undef HWS_BLOCKED_CASE
}
}  // namespace

// Sizes with compile-time blocked kernels (and their sparse variants).
bool fastBlockedSize(int n) { return n >= 256 && n <= 4096 && (n & (n - 1)) == 0; }

// LDS a runtime-radix workgroup may take (the device's 160 KiB per workgroup,
// less room for the kernels' static arrays).
constexpr size_t kGenericLdsBytes = size_t(152) * 1024;

// The blocked layout is used whenever both sizes are multiples of 4 and one
// line of each fits the LDS of one runtime-radix workgroup (n <= 18308); the
// transform plans then also have to fit the workgroup's threads
// (blockedGenericFits: 16 values per thread, n up to about 15000).
bool blockedLayoutSupported(int W, int H) {
  if (fastBlockedSize(W) && fastBlockedSize(H)) return true;
  auto fits = [](int n) { return blockedGenericLds(1, n) <= kGenericLdsBytes; };
  return W % 4 == 0 && H % 4 == 0 && W >= 16 && H >= 16 && fits(W) && fits(H);
}

namespace {
// log2 of the lines (rows or columns) a runtime-radix workgroup of the
// blocked layout transforms: the four of a row block where two such
// workgroups fit the LDS of a CU and the plan the threads, else two, else
// one; -1 if one does not fit.
int genericLinesShift(const FftPlan& plan, int force) {
  auto fits = [&](int lines) {
    return blockedGenericLds(lines, plan.n) <= kGenericLdsBytes &&
           threadsFor(plan, lines) != 0;
  };
  if (force && fits(force)) return force == 4 ? 2 : force == 2 ? 1 : 0;
  // (four lines only while two workgroups fit a CU: at 3072, four lines in
  // one workgroup per CU took 26.6 / 19.5 ms per step for the backward /
  // forward transforms, two lines in three 25.1 / 18.5; one line, whose
  // workgroups read a quarter of each segment, 27.9 / 28.8)
  if (fits(4) && blockedGenericLds(4, plan.n) <= kGenericLdsBytes / 2) return 2;
  if (fits(2)) return 1;
  return fits(1) ? 0 : -1;
}
}  // namespace

bool blockedGenericFits(const hws_gridder* g) {
  return genericLinesShift(g->rowPlan, 0) >= 0 && genericLinesShift(g->colPlan, 0) >= 0;
}

namespace {
size_t colsProjectLds(int H, int nLayersInPass, int tilesY) {
  const size_t words = (size_t(nLayersInPass) * tilesY + 31) / 32;
  const size_t n = size_t(kProjectColumns) * H;
  // transform buffer, n' of every output cell, row bits
  return (n + n / 64) * sizeof(float2) + n * sizeof(double) + words * sizeof(uint32_t);
}
}  // namespace

// The runtime-radix fused kernel: one column per workgroup of at most 512
// threads; its LDS is the transform buffer and the row bits.
size_t colsProjectGenericLds(int H, int nLayersInPass, int tilesY) {
  const size_t words = (size_t(nLayersInPass) * tilesY + 31) / 32;
  return blockedGenericLds(1, H) + words * sizeof(uint32_t);
}

bool colsProjectGeneric(const hws_gridder* g) {
  return !(fastBlockedSize(g->W) && fastBlockedSize(g->H));
}

bool colsProjectSupported(const hws_gridder* g, int nLayersInPass) {
  if (!g->blocked || g->noFusedProject) return false;
  const int T = g->plan.tileSize;
  if (T < 4 || (T & (T - 1)) != 0) return false;
  int maxLds = 65536;
  (void)hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, g->cfg.device);
  if (colsProjectGeneric(g)) {
    const int t = threadsFor(g->colPlan, 1);
    return t != 0 && t <= 512 &&
           colsProjectGenericLds(g->H, nLayersInPass, g->H / T) <= size_t(maxLds);
  }
  return colsProjectLds(g->H, nLayersInPass, g->H / T) <= size_t(maxLds);
}

hws_status launchColsProject(hws_gridder* g, const ColsProjectArgs& a, bool cplx,
                             const float2* layers, float* sum, float* comp, float* isum,
                             float* icomp) {
  if (a.nImg == 0 || a.nLayersInPass == 0) return HWS_OK;
  const bool ref = g->cfg.precision == HWS_PRECISION_REFERENCE;
  if (colsProjectGeneric(g)) {
    const dim3 grid(unsigned(a.W), unsigned(a.nImg));
    const int threads = threadsFor(g->colPlan, 1);
    const size_t lds = colsProjectGenericLds(a.H, a.nLayersInPass, a.tilesY);
    const unsigned long long rad = packRadices(g->colPlan);
    const float2* tw = g->colPlan.twBackward;
#define HWS_COLS_PROJECT_GENERIC(R, C)                                                          \
  hipLaunchKernelGGL((fftColsProjectGenericKernel<R, C>), grid, dim3(threads), lds, g->stream,   \
                     a, layers, rad, tw, sum, comp, isum, icomp)
    if (ref && cplx) HWS_COLS_PROJECT_GENERIC(true, true);
    else if (ref) HWS_COLS_PROJECT_GENERIC(true, false);
    else if (cplx) HWS_COLS_PROJECT_GENERIC(false, true);
    else HWS_COLS_PROJECT_GENERIC(false, false);
#undef HWS_COLS_PROJECT_GENERIC
    return checkHip(g, hipGetLastError(), "fused columns and projection launch");
  }
  const dim3 grid(unsigned(a.W / kProjectColumns), unsigned(a.nImg));
  const size_t lds = colsProjectLds(a.H, a.nLayersInPass, a.tilesY);
  const float2* tw = g->colPlan.twBackward;
  bool ok = false;
  switch (a.H) {
#define HWS_COLS_PROJECT(NN)                                                                    \
  case NN:                                                                                      \
    if (ref && cplx)                                                                            \
      hipLaunchKernelGGL((fftColsProjectKernel<NN, true, true>), grid,                          \
                         dim3(kProjectColumns * NN / 16), lds, g->stream, a, layers, tw, sum,   \
                         comp, isum, icomp);                                                    \
    else if (ref)                                                                               \
      hipLaunchKernelGGL((fftColsProjectKernel<NN, true, false>), grid,                         \
                         dim3(kProjectColumns * NN / 16), lds, g->stream, a, layers, tw, sum,   \
                         comp, isum, icomp);                                                    \
    else if (cplx)                                                                              \
      hipLaunchKernelGGL((fftColsProjectKernel<NN, false, true>), grid,                         \
                         dim3(kProjectColumns * NN / 16), lds, g->stream, a, layers, tw, sum,   \
                         comp, isum, icomp);                                                    \
    else                                                                                        \
      hipLaunchKernelGGL((fftColsProjectKernel<NN, false, false>), grid,                        \
                         dim3(kProjectColumns * NN / 16), lds, g->stream, a, layers, tw, sum,   \
                         comp, isum, icomp);                                                    \
    ok = true;                                                                                  \
    break;
    HWS_COLS_PROJECT(256)
    HWS_COLS_PROJECT(512)
    HWS_COLS_PROJECT(1024)
    HWS_COLS_PROJECT(2048)
    HWS_COLS_PROJECT(4096)
#undef HWS_COLS_PROJECT
    default: break;
  }
  if (!ok) return setLastError(g, HWS_ERR_UNSUPPORTED, "fused projection: height not supported");
  return checkHip(g, hipGetLastError(), "fused columns and projection launch");
}

hws_status fft2d(hws_gridder* g, float2* data, int nbatch, int sign, const FftSparse* sparse,
                 const ModelSource* model) {
  if (nbatch <= 0) return HWS_OK;
  const int W = g->W, H = g->H;
  if (g->blocked && !(fastBlockedSize(W) && fastBlockedSize(H))) {
    // blocked layout, runtime radices (sparse modes as for the compile-time
    // kernels)
    const float2* twRow = sign > 0 ? g->rowPlan.twBackward : g->rowPlan.twForward;
    const float2* twCol = sign > 0 ? g->colPlan.twBackward : g->colPlan.twForward;
    // four rows (columns) per workgroup where they fit the LDS, else two, else one
    const int rbShift = genericLinesShift(g->rowPlan, g->fftLines);
    const int cbShift = genericLinesShift(g->colPlan, g->fftLines);
    if (rbShift < 0 || cbShift < 0)
      return setLastError(g, HWS_ERR_UNSUPPORTED, "transform does not fit one workgroup");
    const int tr = threadsFor(g->rowPlan, 1 << rbShift), tc = threadsFor(g->colPlan, 1 << cbShift);
    const int rl = 1 << rbShift, cl = 1 << cbShift;
    const size_t ldsR = blockedGenericLds(rl, W), ldsC = blockedGenericLds(cl, H);
    const unsigned long long radR = packRadices(g->rowPlan), radC = packRadices(g->colPlan);
    const int nColBlocks = W >> cbShift;
    GenericSparseArgs rs{}, cs{};
    ModelSource msrc{};
    if (sparse) {
      rs.layer0 = cs.layer0 = sparse->layer0;
      rs.tilesY = cs.tilesY = sparse->tilesY;
      rs.tshift = cs.tshift = sparse->tshift;
      rs.tilesX = cs.tilesX = W >> sparse->tshift;
      if (sign > 0) {
        rs.bits = sparse->occ;
        cs.bits = sparse->rowLive;
      } else {
        cs.bits = sparse->occ;  // the need bits
      }
    }
    if (model) msrc = *model;
    const dim3 gr(H >> rbShift, nbatch), gc(nColBlocks, nbatch);
    hipStream_t st = g->stream;
    if (sign > 0) {
      if (sparse) {
        launchRowsGeneric<+1, 1, false>(tr, rl, gr, ldsR, st, data, W, H, radR, twRow, rs, msrc);
        // (rowsOnly: the columns pass runs fused with the projection)
        if (!sparse->rowsOnly)
          launchColsGeneric<+1, 1>(tc, cl, gc, ldsC, st, data, W, H, nColBlocks, radC, twCol, cs);
      } else {
        launchRowsGeneric<+1, 0, false>(tr, rl, gr, ldsR, st, data, W, H, radR, twRow, rs, msrc);
        launchColsGeneric<+1, 0>(tc, cl, gc, ldsC, st, data, W, H, nColBlocks, radC, twCol, cs);
      }
    } else {
      // (the mixed precision: the reference precision's FP64 sincos leaves no
      // registers for the cells it keeps)
      if (model && sparse && sparse->colNeed && model->precision != HWS_PRECISION_REFERENCE)
        launchModelLayersGeneric(tr, rl, dim3(H >> rbShift), ldsR, st, data, W, H, nbatch, radR,
                                 twRow, msrc, sparse->colNeed, sparse->tshift);
      else if (model && model->precision == HWS_PRECISION_REFERENCE)
        launchRowsGeneric<-1, 2, true>(tr, rl, gr, ldsR, st, data, W, H, radR, twRow, rs, msrc);
      else if (model)
        launchRowsGeneric<-1, 2, false>(tr, rl, gr, ldsR, st, data, W, H, radR, twRow, rs, msrc);
      else
        launchRowsGeneric<-1, 0, false>(tr, rl, gr, ldsR, st, data, W, H, radR, twRow, rs, msrc);
      if (sparse)
        launchColsGeneric<-1, 2>(tc, cl, gc, ldsC, st, data, W, H, nColBlocks, radC, twCol, cs);
      else
        launchColsGeneric<-1, 0>(tc, cl, gc, ldsC, st, data, W, H, nColBlocks, radC, twCol, cs);
    }
    return checkHip(g, hipGetLastError(), "fft blocked generic launch");
  }
  if (g->blocked) {
    const float2* twRow = sign > 0 ? g->rowPlan.twBackward : g->rowPlan.twForward;
    const float2* twCol = sign > 0 ? g->colPlan.twBackward : g->colPlan.twForward;
    bool ok;
    if (sign > 0 && sparse) {
      ok = false;
      switch (W) {
#define HWS_SPARSE_CASE(NN)                                                                    \
  case NN:                                                                                     \
    hipLaunchKernelGGL((fftRowsBlockedSparseKernel<+1, NN>), dim3(H / 4, nbatch), dim3(NN / 4), \
                       0, g->stream, data, H, twRow, sparse->occ, sparse->layer0,              \
                       sparse->tilesY, sparse->tshift);                                        \
    ok = true;                                                                                 \
    break;
        HWS_SPARSE_CASE(256)
        HWS_SPARSE_CASE(512)
        HWS_SPARSE_CASE(1024)
        HWS_SPARSE_CASE(2048)
        HWS_SPARSE_CASE(4096)
#undef HWS_SPARSE_CASE
        default: break;
      }
      bool cols = sparse->rowsOnly;
      if (!cols) switch (H) {
#define HWS_ROWSPARSE_COLS(NN)                                                                 \
  case NN:                                                                                     \
    hipLaunchKernelGGL((fftColsBlockedRowSparseKernel<+1, NN>), dim3(W / 4, nbatch),            \
                       dim3(NN / 4), 0, g->stream, data, W, twCol, sparse->rowLive,            \
                       sparse->layer0, sparse->tilesY, sparse->tshift);                        \
    cols = true;                                                                               \
    break;
        HWS_ROWSPARSE_COLS(256)
        HWS_ROWSPARSE_COLS(512)
        HWS_ROWSPARSE_COLS(1024)
        HWS_ROWSPARSE_COLS(2048)
        HWS_ROWSPARSE_COLS(4096)
#undef HWS_ROWSPARSE_COLS
        default: break;
      }
      ok = ok && cols;
    } else if (sign < 0 && sparse) {
      if (model && sparse->colNeed) {
        // every layer per workgroup, the model and n' read and computed once
        ok = false;
        const bool ref = model->precision == HWS_PRECISION_REFERENCE;
        const bool cplx = model->cplx != 0;
        const dim3 mgrid(H / kProjectColumns);
        switch (W) {
#define HWS_MODEL_LAYERS(NN)                                                                   \
  case NN:                                                                                     \
    if (ref && cplx)                                                                           \
      hipLaunchKernelGGL((fftRowsModelLayersKernel<NN, true, true>), mgrid,                    \
                         dim3(kProjectColumns * NN / 16), 0, g->stream, data, H, nbatch,        \
                         twRow, *model, sparse->colNeed, sparse->tshift);                       \
    else if (ref)                                                                              \
      hipLaunchKernelGGL((fftRowsModelLayersKernel<NN, true, false>), mgrid,                   \
                         dim3(kProjectColumns * NN / 16), 0, g->stream, data, H, nbatch,        \
                         twRow, *model, sparse->colNeed, sparse->tshift);                       \
    else if (cplx)                                                                             \
      hipLaunchKernelGGL((fftRowsModelLayersKernel<NN, false, true>), mgrid,                   \
                         dim3(kProjectColumns * NN / 16), 0, g->stream, data, H, nbatch,        \
                         twRow, *model, sparse->colNeed, sparse->tshift);                       \
    else                                                                                       \
      hipLaunchKernelGGL((fftRowsModelLayersKernel<NN, false, false>), mgrid,                  \
                         dim3(kProjectColumns * NN / 16), 0, g->stream, data, H, nbatch,        \
                         twRow, *model, sparse->colNeed, sparse->tshift);                       \
    ok = true;                                                                                 \
    break;
          HWS_MODEL_LAYERS(256)
          HWS_MODEL_LAYERS(512)
          HWS_MODEL_LAYERS(1024)
          HWS_MODEL_LAYERS(2048)
          HWS_MODEL_LAYERS(4096)
#undef HWS_MODEL_LAYERS
          default: break;
        }
      } else if (model) {
        ok = false;
        const int ml = model->precision == HWS_PRECISION_REFERENCE ? modelLayers<true>()
                                                                   : modelLayers<false>();
        const dim3 mgrid(H / 4, (nbatch + ml - 1) / ml);
        switch (W) {
#define HWS_MODEL_ROWS(NN)                                                                      \
  case NN:                                                                                      \
    if (model->precision == HWS_PRECISION_REFERENCE)                                            \
      hipLaunchKernelGGL((fftRowsFromModelKernel<-1, NN, true>), mgrid, dim3(NN / 4), 0,         \
                         g->stream, data, H, nbatch, twRow, *model);                            \
    else                                                                                        \
      hipLaunchKernelGGL((fftRowsFromModelKernel<-1, NN, false>), mgrid, dim3(NN / 4), 0,        \
                         g->stream, data, H, nbatch, twRow, *model);                            \
    ok = true;                                                                                  \
    break;
          HWS_MODEL_ROWS(256)
          HWS_MODEL_ROWS(512)
          HWS_MODEL_ROWS(1024)
          HWS_MODEL_ROWS(2048)
          HWS_MODEL_ROWS(4096)
#undef HWS_MODEL_ROWS
          default: break;
        }
      } else {
        ok = launchBlocked<-1>(W, true, data, W, H, nbatch, twRow, g->stream);
      }
      bool cols = false;
      switch (H) {
#define HWS_SPARSE_COLS(NN)                                                                    \
  case NN:                                                                                     \
    hipLaunchKernelGGL((fftColsBlockedSparseKernel<-1, NN>), dim3(W / 4, nbatch), dim3(NN / 4), \
                       0, g->stream, data, W, twCol, sparse->occ, sparse->layer0,              \
                       sparse->tilesY, sparse->tshift);                                        \
    cols = true;                                                                               \
    break;
        HWS_SPARSE_COLS(256)
        HWS_SPARSE_COLS(512)
        HWS_SPARSE_COLS(1024)
        HWS_SPARSE_COLS(2048)
        HWS_SPARSE_COLS(4096)
#undef HWS_SPARSE_COLS
        default: break;
      }
      ok = ok && cols;
    } else if (sign > 0)
      ok = launchBlocked<+1>(W, true, data, W, H, nbatch, twRow, g->stream) &&
           launchBlocked<+1>(H, false, data, W, H, nbatch, twCol, g->stream);
    else
      ok = launchBlocked<-1>(W, true, data, W, H, nbatch, twRow, g->stream) &&
           launchBlocked<-1>(H, false, data, W, H, nbatch, twCol, g->stream);
    if (!ok) return setLastError(g, HWS_ERR_UNSUPPORTED, "blocked layout size not supported");
    return checkHip(g, hipGetLastError(), "fft blocked launch");
  }
  int maxLds = 65536;
  (void)hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, g->cfg.device);
  const float2* twRow = sign > 0 ? g->rowPlan.twBackward : g->rowPlan.twForward;
  const float2* twCol = sign > 0 ? g->colPlan.twBackward : g->colPlan.twForward;
  bool rowsDone = false, colsDone = false;
  if (sign > 0) {
    rowsDone = dispatchPow2<+1>(W, data, W, H, nbatch, twRow, twCol, true, g->stream);
    colsDone = dispatchPow2<+1>(H, data, W, H, nbatch, twRow, twCol, false, g->stream);
  } else {
    rowsDone = dispatchPow2<-1>(W, data, W, H, nbatch, twRow, twCol, true, g->stream);
    colsDone = dispatchPow2<-1>(H, data, W, H, nbatch, twRow, twCol, false, g->stream);
  }
  if (rowsDone || colsDone) {
    hws_status st = checkHip(g, hipGetLastError(), "fft pow2 launch");
    if (st != HWS_OK) return st;
  }
  // rows: as many whole rows per block as fit 4096 values (at least one)
  if (!rowsDone) {
    int rows = std::max(1, 4096 / W);
    rows = std::min(rows, H);
    const int threads = threadsFor(g->rowPlan, rows);
    const size_t lds = size_t(rows) * ldsPitch(W) * sizeof(float2);
    if (threads == 0 || lds > size_t(maxLds))
      return setLastError(g, HWS_ERR_UNSUPPORTED, "row transform does not fit one workgroup");
    dim3 grid((H + rows - 1) / rows, nbatch);
    const unsigned long long radices = packRadices(g->rowPlan);
    const float2* tw = sign > 0 ? g->rowPlan.twBackward : g->rowPlan.twForward;
    if (sign > 0)
      launchRowsPlain<+1>(threads, grid, lds, g->stream, data, W, H, rows, radices, tw);
    else
      launchRowsPlain<-1>(threads, grid, lds, g->stream, data, W, H, rows, radices, tw);
    hws_status st = checkHip(g, hipGetLastError(), "fft rows launch");
    if (st != HWS_OK) return st;
  }
  // columns: blocks of bc columns, [bc][ldsPitch(H)] in LDS
  if (!colsDone) {
    int bc = std::max(1, 8192 / H);
    bc = std::min(std::min(bc, 16), W);
    while (bc > 1 && size_t(bc) * ldsPitch(H) * sizeof(float2) > size_t(maxLds)) --bc;
    int threads = threadsFor(g->colPlan, bc);
    while (threads == 0 && bc > 1) threads = threadsFor(g->colPlan, --bc);
    const size_t lds = size_t(bc) * size_t(ldsPitch(H)) * sizeof(float2);
    if (threads == 0 || lds > size_t(maxLds))
      return setLastError(g, HWS_ERR_UNSUPPORTED, "column transform does not fit one workgroup");
    const int nColBlocks = (W + bc - 1) / bc;
    dim3 grid(nColBlocks, nbatch);
    const unsigned long long radices = packRadices(g->colPlan);
    const float2* tw = sign > 0 ? g->colPlan.twBackward : g->colPlan.twForward;
    if (sign > 0)
      launchColsPlain<+1>(threads, grid, lds, g->stream, data, W, H, bc, nColBlocks, radices, tw);
    else
      launchColsPlain<-1>(threads, grid, lds, g->stream, data, W, H, bc, nColBlocks, radices, tw);
    hws_status st = checkHip(g, hipGetLastError(), "fft cols launch");
    if (st != HWS_OK) return st;
  }
  return HWS_OK;
}

}  // namespace hws
