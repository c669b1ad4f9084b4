// fft_core.hpp -- register/LDS Stockham FFT building blocks for CDNA4 (gfx950).
//
// A block transforms a batch of pencils of one length N.  Each thread owns
// a slice t in [0, P) of G pencils; a stage of radix R does the N/R
// butterflies of every pencil, butterfly j = t + P*m going to thread t.
// Stage s reads its R inputs at j + r*N/R, twiddles them by
// exp(sign*2*pi*i*r*k/(Ns*R)) with k = j mod Ns (Ns = product of the
// earlier radices), does a radix-R DFT in registers and writes the outputs
// to (j/Ns)*Ns*R + k + r*Ns (the self-sorting Stockham form, Govindaraju
// et al., SC'08).  The first stage takes its inputs from registers filled
// by the caller (straight from HBM) and the last stage leaves its outputs
// in registers (written by the caller straight to HBM), so a transform of
// S stages makes S-1 round trips through LDS.  The same positions
// j + r*N/R are used for the first-stage inputs and (with the last radix)
// for the last-stage outputs, so callers know which elements each thread
// holds.
//
// Twiddles come from a host-built table (long-double sin/cos rounded once);
// stage s uses entries [off(s) + k*(R-1) + (r-1)] = exp(-2*pi*i*r*k/(Ns*R)),
// conjugated for the inverse transform.  FP64 transforms read only the r = 1
// entry and form the powers by multiplication.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gt {

// f(std::integral_constant<int, I>) for I = B .. E-1, unrolled at compile time
template <int B, int E, class F>
__device__ __forceinline__ void static_for_fft(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for_fft<B + 1, E>(f);
  }
}

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// a * (S*i), S = +1 or -1
template <int S, typename T>
__device__ __forceinline__ cx<T> mul_si(cx<T> a) {
  if constexpr (S > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }

// ---------------------------------------------------------------------------
// Radix plans.  Radices are applied in the listed order.
// ---------------------------------------------------------------------------
struct PlanInfo {
  int nst;
  int r[6];
};

__host__ __device__ constexpr PlanInfo plan_of(int n) {
  switch (n) {
    case 8: return {1, {8, 0, 0, 0, 0, 0}};
    case 16: return {2, {4, 4, 0, 0, 0, 0}};
    case 24: return {2, {3, 8, 0, 0, 0, 0}};
    case 32: return {2, {4, 8, 0, 0, 0, 0}};
    case 48: return {3, {3, 2, 8, 0, 0, 0}};
    case 64: return {2, {8, 8, 0, 0, 0, 0}};
    case 96: return {3, {3, 4, 8, 0, 0, 0}};
    case 128: return {3, {8, 2, 8, 0, 0, 0}};  // first radix = last (three-wave x pass, fused z passes)
    case 192: return {3, {3, 8, 8, 0, 0, 0}};
    case 256: return {3, {8, 4, 8, 0, 0, 0}};  // first radix = last (three-wave x pass, fused z passes)
    case 384: return {4, {3, 2, 8, 8, 0, 0}};
    case 512: return {3, {8, 8, 8, 0, 0, 0}};
    case 768: return {4, {3, 4, 8, 8, 0, 0}};
    case 1024: return {4, {2, 8, 8, 8, 0, 0}};
    default: return {0, {0, 0, 0, 0, 0, 0}};
  }
}

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ constexpr bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <int N>
struct Plan {
  static constexpr PlanInfo I = plan_of(N);
  static_assert(I.nst > 0, "unsupported FFT length");
  static constexpr int S = I.nst;
  static constexpr int radix(int s) { return I.r[s]; }
  static constexpr int ns(int s) {
    int v = 1;
    for (int q = 0; q < s; q++) v *= I.r[q];
    return v;
  }
  // offset of stage s in the twiddle table
  static constexpr int twoff(int s) {
    int o = 0;
    for (int q = 0; q < s; q++) o += ns(q) * (I.r[q] - 1);
    return o;
  }
  static constexpr int twsize() { return twoff(S); }
};

// ---------------------------------------------------------------------------
// Radix-R DFTs in registers, y[q] = sum_r v[r] exp(S*2*pi*i*r*q/R).
// ---------------------------------------------------------------------------
template <typename T>
struct K {
  static constexpr T h2 = T(0.70710678118654752440084436210484904);  // sqrt(2)/2
  static constexpr T h3 = T(0.86602540378443864676372317075293618);  // sqrt(3)/2
};

template <int S, typename T>
__device__ __forceinline__ void dft2(cx<T>* v) {
  cx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S, typename T>
__device__ __forceinline__ void dft3(cx<T>* v) {
  cx<T> a = v[0], b = v[1], c = v[2];
  cx<T> s = cadd(b, c);
  cx<T> d = csub(b, c);
  v[0] = cadd(a, s);
  cx<T> m = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
  cx<T> e = mul_si<S>(cscale(d, K<T>::h3));
  v[1] = cadd(m, e);
  v[2] = csub(m, e);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cx<T>* v) {
  cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  cx<T> c = cadd(v[1], v[3]), d = mul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

template <int S, typename T>
__device__ __forceinline__ void dft8(cx<T>* v) {
  // decimation in frequency: a[n] = v[n] + v[n+4], b[n] = (v[n] - v[n+4]) w^n
  cx<T> a[4], b[4];
#pragma unroll
  for (int n = 0; n < 4; n++) {
    a[n] = cadd(v[n], v[n + 4]);
    b[n] = csub(v[n], v[n + 4]);
  }
  const T h = K<T>::h2;
  // w = (1 + S i)/sqrt2 ; w^2 = S i ; w^3 = (-1 + S i)/sqrt2
  b[1] = {(b[1].x - T(S) * b[1].y) * h, (b[1].y + T(S) * b[1].x) * h};
  b[2] = mul_si<S>(b[2]);
  b[3] = {(-b[3].x - T(S) * b[3].y) * h, (-b[3].y + T(S) * b[3].x) * h};
  dft4<S>(a);
  dft4<S>(b);
  v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
  v[1] = b[0]; v[3] = b[1]; v[5] = b[2]; v[7] = b[3];
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(cx<T>* v) {
  if constexpr (R == 2) dft2<S>(v);
  else if constexpr (R == 3) dft3<S>(v);
  else if constexpr (R == 4) dft4<S>(v);
  else if constexpr (R == 8) dft8<S>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// LDS layouts.  off(g, n) is the complex offset of element n of the g-th
// pencil the thread works on (p = pencil index within the block).
// ---------------------------------------------------------------------------

// LDS banks on gfx950 (measured with the SQ_LDS_BANK_CONFLICT counter): a
// 128-bit write is served 8 lanes at a time and is conflict-free when those
// lanes hit distinct 16-byte slots of a 128-byte window; a 128-bit read, 16
// lanes at a time over 256 bytes; a 64-bit write, 16 lanes over 128 bytes;
// a 64-bit read, 32 lanes over 256 bytes.
//
// T pencils interleaved: element n of pencil p at n'*T + p.  With 16-byte
// elements, or 16 pencils of 8 bytes, n' = n and every stage exchange is
// conflict-free.  With 8 pencils of 8 bytes (SW), a 64-bit write's 16 lanes
// span two slices t whose stage-0 outputs n = 8t + r differ only in bit 3,
// so n' = n ^ ((n >> 3) & 1) puts them in different halves of the window
// (a permutation of [0, N) for even N).
template <int T, bool SW = false>
struct LayInterleaved {
  static constexpr bool XOR_LINEAR = false;
  int p;
  __device__ __forceinline__ int off(int /*g*/, int n) const {
    if constexpr (SW) return (n ^ ((n >> 3) & 1)) * T + p;
    else return n * T + p;
  }
  __host__ __device__ static constexpr int swz(int n) { return n; }
  __device__ __forceinline__ int row(int /*g*/) const { return 0; }
};

// Pencils stored one after another with one pad word every 8 (stride
// N + N/8): used when lanes run along n (contiguous rows in HBM).  The pad
// spreads the stride-R writes of the early Stockham stages over the banks.
template <int N>
struct LayRows {
  static constexpr bool XOR_LINEAR = false;
  int s0;     // first pencil of the thread
  int sstep;  // pencil step between the thread's G pencils
  static constexpr int STR = N + N / 8;
  __device__ __forceinline__ int off(int g, int n) const { return (s0 + g * sstep) * STR + n + (n >> 3); }
  __host__ __device__ static constexpr int swz(int n) { return n; }
  __device__ __forceinline__ int row(int /*g*/) const { return 0; }
};

// Rows one after another, element n of a row at swz(n), with no pad words.
// For a power-of-two N, swz(n) = n ^ ((n >> 3) & 15): with the serving
// widths above, the stage exchanges of the radix-8 stages (writes to
// 8j + r and to 64(j/8) + j%8 + 8r, reads from j + 64r for lane j) hit
// distinct slots in both precisions.  Other lengths keep n ^ ((n >> 4) & 15),
// which permutes each aligned block of 16 elements and, for lengths with a
// partial last block of 8 (24, 48), stays inside that block.
//
// The swizzle is linear over XOR: when a and c have no bit in common,
// swz(a | c) = swz(a) ^ swz(c).  A stage position a + c with a the thread's
// part and c a compile-time multiple of a power of two is such a sum, so for
// a power-of-two N the byte address of every element of a stage exchange is
// one per-thread base XOR a constant (XOR_LINEAR), instead of a swizzle per
// element.
template <int N>
struct LaySwz {
  static constexpr bool XOR_LINEAR = (N & (N - 1)) == 0;
  int s0;     // first row of the thread
  int sstep;  // row step between the thread's G rows
  __host__ __device__ static constexpr int swz(int n) {
    return ((N & (N - 1)) == 0) ? (n ^ ((n >> 3) & 15)) : (n ^ ((n >> 4) & 15));
  }
  __device__ __forceinline__ int row(int g) const { return (s0 + g * sstep) * N; }
  __device__ __forceinline__ int off(int g, int n) const { return row(g) + swz(n); }
};

// Synchronisation between the stages of a transform: the whole block, or
// only the wavefront when every pencil of the transform lives in one wave.
enum { SYNC_BLOCK = 0, SYNC_WAVE = 1 };

template <int SYNC>
__device__ __forceinline__ void stage_sync() {
  if constexpr (SYNC == SYNC_BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// FP64 twiddles held in registers: the r = 1 table entry of butterfly
// j = t + P*m of every stage s >= 1, at slot slot(s) + m.  A kernel that runs
// several transforms of one length (inverse or forward: both read the same
// entries) loads them once, with its data, so no stage waits for a table
// load.
template <typename T, int N, int P>
struct TwRegs {
  using PL = Plan<N>;
  static constexpr int slot(int s) {
    int v = 0;
    for (int q = 1; q < s; q++) v += cdiv(N / PL::radix(q), P);
    return v;
  }
  static constexpr int NW = slot(PL::S);
  cx<T> w[NW > 0 ? NW : 1];
  __device__ __forceinline__ void load(const cx<T>* __restrict__ tw, int t) {
    static_assert(sizeof(T) == 8, "FP64 transforms form the powers of the r = 1 entry");
    static_for_fft<1, PL::S>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      constexpr int Rs = PL::radix(s), Ns = PL::ns(s), NR = N / Rs, M = cdiv(NR, P);
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int j = t + P * m;
        // a thread past the last butterfly reads entry 0 and never uses it
        const int k = (NR % P == 0 || j < NR) ? j % Ns : 0;
        w[slot(s) + m] = tw[PL::twoff(s) + k * (Rs - 1)];
      }
    });
  }
};

// Copy the twiddle table of length N (Plan<N>::twsize() entries) into LDS,
// cooperatively over nthr threads; a transform then reads its twiddles with
// LDS latency instead of L2 latency (run() with the LDS copy as its table).
// The copy is visible to the other wavefronts after the next block barrier.
template <typename T, int N>
__device__ __forceinline__ void tw_to_lds(const cx<T>* __restrict__ tw, cx<T>* twl, int tid, int nthr) {
  constexpr int TS = Plan<N>::twsize();
  for (int i = tid; i < TS; i += nthr) twl[i] = tw[i];
}

// ---------------------------------------------------------------------------
// The staged transform.
// ---------------------------------------------------------------------------
template <typename T, int N, int P, int G, int SIGN, class L, int SYNC = SYNC_BLOCK>
struct Fft {
  using PL = Plan<N>;
  static constexpr int S = PL::S;
  static constexpr int R0 = PL::radix(0);
  static constexpr int RL = PL::radix(S - 1);
  static constexpr int NR0 = N / R0;
  static constexpr int NRL = N / RL;
  static constexpr int M0 = cdiv(NR0, P);
  static constexpr int ML = cdiv(NRL, P);
  static constexpr int E0 = M0 * R0;  // first-stage values per pencil per thread
  static constexpr int EL = ML * RL;  // last-stage values per pencil per thread

  // element index held in slot (m, r) of the first / last stage
  __device__ static __forceinline__ int in_pos(int t, int m, int r) { return t + P * m + r * NR0; }
  __device__ static __forceinline__ bool in_ok(int t, int m) { return (NR0 % P == 0) || (t + P * m < NR0); }
  __device__ static __forceinline__ int out_pos(int t, int m, int r) { return t + P * m + r * NRL; }
  __device__ static __forceinline__ bool out_ok(int t, int m) { return (NRL % P == 0) || (t + P * m < NRL); }

  using Twr = TwRegs<T, N, P>;
  static constexpr int twn(int s) { return Twr::slot(s); }
  __device__ static __forceinline__ void load_tw(const cx<T>* __restrict__ tw, int t, Twr& R) { R.load(tw, t); }

  template <int s, bool PRE = false>
  __device__ static __forceinline__ void stage(const L& lay, cx<T>* lds, const cx<T>* __restrict__ tw, int t,
                                               cx<T> (&in)[G][E0], cx<T> (&out)[G][EL], const Twr* twr = nullptr) {
    constexpr int R = PL::radix(s);
    constexpr int Ns = PL::ns(s);
    constexpr int NR = N / R;
    constexpr int M = cdiv(NR, P);
    constexpr bool FULL = (NR % P == 0);
    cx<T> v[G][M * R];
    constexpr int EB = (int)sizeof(cx<T>);
    if constexpr (s == 0) {
#pragma unroll
      for (int g = 0; g < G; g++)
#pragma unroll
        for (int q = 0; q < M * R; q++) v[g][q] = in[g][q];
    } else {
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int j = t + P * m;
        if (FULL || j < NR) {
#pragma unroll
          for (int g = 0; g < G; g++) {
            if constexpr (L::XOR_LINEAR && pow2(NR)) {
              // j < NR and r*NR share no bit: one base, one XOR per element
              const int b = (lay.row(g) + L::swz(j)) * EB;
              const char* base = reinterpret_cast<const char*>(lds);
#pragma unroll
              for (int r = 0; r < R; r++)
                v[g][m * R + r] = *reinterpret_cast<const cx<T>*>(base + (b ^ (L::swz(r * NR) * EB)));
            } else {
#pragma unroll
              for (int r = 0; r < R; r++) v[g][m * R + r] = lds[lay.off(g, j + r * NR)];
            }
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
      const int j = t + P * m;
      if (FULL || j < NR) {
        if constexpr (s > 0) {
          const int k = j % Ns;
          const cx<T>* w = tw + PL::twoff(s) + k * (R - 1);
          if constexpr (sizeof(T) == 8) {
            // FP64: w^r = w^(r-1) w from the first table entry, applied as it
            // is formed, so two twiddles are live instead of R-1 (register
            // pressure sets the occupancy of the fused x pass); the products
            // add at most R-2 roundings to w^r
            cx<T> w1;
            if constexpr (PRE) {
              w1 = twr->w[twn(s) + m];
              // an empty asm makes w1 a value of this point of the stage, so
              // that the powers formed from it are not hoisted to the kernel
              // start and held in registers through every earlier stage
              asm volatile("" : "+v"(w1.x), "+v"(w1.y));
            } else {
              w1 = w[0];
            }
            cx<T> wk = w1;
#pragma unroll
            for (int r = 1; r < R; r++) {
              if (r > 1) wk = cmul(wk, w1);
#pragma unroll
              for (int g = 0; g < G; g++) {
                if constexpr (SIGN < 0) v[g][m * R + r] = cmul(v[g][m * R + r], wk);
                else v[g][m * R + r] = cmulc(v[g][m * R + r], wk);
              }
            }
          } else {
            cx<T> wr[R - 1];
#pragma unroll
            for (int r = 1; r < R; r++) wr[r - 1] = w[r - 1];
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
              for (int r = 1; r < R; r++) {
                if constexpr (SIGN < 0) v[g][m * R + r] = cmul(v[g][m * R + r], wr[r - 1]);
                else v[g][m * R + r] = cmulc(v[g][m * R + r], wr[r - 1]);
              }
          }
        }
#pragma unroll
        for (int g = 0; g < G; g++) dft<R, SIGN>(&v[g][m * R]);
      }
    }
    if constexpr (s == S - 1) {
#pragma unroll
      for (int g = 0; g < G; g++)
#pragma unroll
        for (int q = 0; q < M * R; q++) out[g][q] = v[g][q];
    } else {
      stage_sync<SYNC>();  // every read of the previous stage (or previous use) is done
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int j = t + P * m;
        if (FULL || j < NR) {
          const int k = j % Ns;
          const int d = (j / Ns) * Ns * R + k;
#pragma unroll
          for (int g = 0; g < G; g++) {
            if constexpr (L::XOR_LINEAR && pow2(Ns) && pow2(R)) {
              // k < Ns, r*Ns < Ns*R and the multiple of Ns*R share no bit
              const int b = (lay.row(g) + L::swz(d)) * EB;
              char* base = reinterpret_cast<char*>(lds);
#pragma unroll
              for (int r = 0; r < R; r++)
                *reinterpret_cast<cx<T>*>(base + (b ^ (L::swz(r * Ns) * EB))) = v[g][m * R + r];
            } else {
#pragma unroll
              for (int r = 0; r < R; r++) lds[lay.off(g, d + r * Ns)] = v[g][m * R + r];
            }
          }
        }
      }
      stage_sync<SYNC>();
      stage<s + 1, PRE>(lay, lds, tw, t, in, out, twr);
    }
  }

  // Transform the values in `in` (first-stage layout) into `out` (last-stage
  // layout).  All threads of the block must call it (it has barriers).
  __device__ static __forceinline__ void run(const L& lay, cx<T>* lds, const cx<T>* __restrict__ tw, int t,
                                             cx<T> (&in)[G][E0], cx<T> (&out)[G][EL]) {
    stage<0>(lay, lds, tw, t, in, out);
  }
  // The same with the twiddles of load_tw (FP64).
  __device__ static __forceinline__ void run(const L& lay, cx<T>* lds, const cx<T>* __restrict__ tw, const Twr& twr,
                                             int t, cx<T> (&in)[G][E0], cx<T> (&out)[G][EL]) {
    stage<0, true>(lay, lds, tw, t, in, out, &twr);
  }
};

}  // namespace gt
