// gt_device.hpp -- device helpers shared by the step kernels and the
// adjoint kernels.
#pragma once
#include "gt_internal.hpp"

#include <type_traits>

namespace gt {

// The transform kernels take their LDS from the dynamic segment, sized by the
// launch (lds_bytes), rather than from a static __shared__ array.  Measured on
// MI355X with 384-thread blocks of 48 KB: a static allocation averages 1.96
// resident blocks per CU in steady state (3 fit), a dynamic one 2.75, and the
// x pass runs 6% faster for it.
#define GT_DYN_LDS(T, name)                                  \
  extern __shared__ __align__(16) unsigned char name##_raw[]; \
  T* name = reinterpret_cast<T*>(name##_raw)

// FP32 transforms read their twiddles (R-1 per butterfly and stage) from a
// copy of the table in LDS rather than from the global table (tw_to_lds)
#ifndef GT_TWLDS32
#define GT_TWLDS32 1
#endif

// f(std::integral_constant<int, I>) for I = B .. E-1, unrolled at compile time
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// integrating factor exp(-0.5*dt*(nu*k2**alpnu + mu*k2**alpmu)), 1 at k = 0
// (src/defk.f90:57-68), evaluated on the fly instead of read from a table
__device__ __forceinline__ double emk_of(const KP& kp, double fk2) {
  double scra = 0.0;
  if (fk2 != 0.0) {
    if (kp.visc_simple) scra = kp.nu * fk2;
    else scra = kp.nu * pow(fk2, kp.alpnu) + kp.mu * pow(fk2, kp.alpmu);
  }
  double e = exp(-0.5 * kp.dt * scra);
  if (fk2 == 0.0) e = 1.0;
  return e;
}

// integrating factor at mode (kx, ky, kz) with k^2 = fk2.  With simple
// viscosity (alpnu = 1, mu = 0) exp(-0.5*dt*nu*k2) is the product of one
// factor per axis, so three small tables replace the exponential (a few
// roundings from it, and exactly 1 at k = 0); otherwise emk_of.
__device__ __forceinline__ double emk_at(const KP& kp, int kx, int ky, int kz, double fk2) {
  if (kp.emx) return (kp.emx[kx] * kp.emy[ky]) * kp.emz[kz];
  return emk_of(kp, fk2);
}

// Whether any thread of the wavefront has alive = true.  In the pencil
// kernels every wavefront holds all T pencils of its block (pencil = lane
// mod T), so this is the block's answer too, with no LDS word and no wait.
__device__ __forceinline__ bool wave_any(bool alive) { return __any(alive ? 1 : 0) != 0; }

// mode (kx, ky, kz) is zeroed by trunc: the ellipsoidal test with the sums in
// the reference order, and k = 0 (src/trunc.f90:14-19,27); never when
// truncate /= 1
__device__ __forceinline__ bool trunc_zero(const KP& kp, int kx, int ky, int kz) {
  if (kp.truncate != 1) return false;
  if (kx == 0 && ky == 0 && kz == 0) return true;
  return (kp.fkzn[kz] + kp.fkyn[ky]) + kp.fkxn[kx] > kp.rtrunc2;
}

// z = row0 + i*row1 packing of two half-spectrum rows for one complex
// transform of length N (conj_half: element above N/2)
template <typename W>
__device__ __forceinline__ cx<W> herm_pack(cx<W> X0, cx<W> X1, bool conj_half) {
  if (conj_half) return {X0.x + X1.y, -X0.y + X1.x};
  return {X0.x - X1.y, X0.y + X1.x};
}

// Element i of a row whose start is the same for the whole wavefront, with
// the offset formed in 32 bits (a scalar base plus an unsigned vector
// offset, instead of a 64-bit address per lane); i * sizeof < 2^32.
template <typename V>
__device__ __forceinline__ V ld_row(const V* row, unsigned i) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(row) + i * (unsigned)sizeof(V));
}
template <typename V>
__device__ __forceinline__ void st_row(V* row, unsigned i, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(row) + i * (unsigned)sizeof(V)) = v;
}
// Entry p (0 <= p < T) of T consecutive int16 extents that start at a
// block-uniform, 4-byte aligned address.  They are read through the constant
// address space, so the loads are scalar and hit the small per-CU-group
// constant cache; a vector load instead would be one full memory latency
// that every dependent load of the block waits behind.
template <int T>
__device__ __forceinline__ int tile_ext(const int16_t* e, int p) {
  static_assert(T % 2 == 0, "pairs of entries per 32-bit word");
  typedef __attribute__((address_space(4))) const int CInt;
  CInt* e32 = (CInt*)(e);
  int v = -1;
#pragma unroll
  for (int i = 0; i < T / 2; i++) {
    const int w = e32[i];
    const int lo = (int)(int16_t)(w & 0xffff), hi = w >> 16;  // entries 2i, 2i+1
    v = (p == 2 * i) ? lo : v;
    v = (p == 2 * i + 1) ? hi : v;
  }
  return v;
}

// A thread slot of a transform of length N with P threads per sequence
// covers positions t + C, 0 <= t < P.  With every mode index min(n, N - n)
// of the spectral side at most a cut below 3N/8, a slot whose positions all
// lie in [3N/8, 5N/8) holds only zeros on input (inverse transforms) or
// values that are never stored (forward transforms): it is pruned at
// compile time, so the butterflies fold the zeros or drop the dead outputs.
template <int N, int P, int C>
__host__ __device__ constexpr bool mid_slot() {
  return 8 * C >= 3 * N && 8 * (C + P) <= 5 * N;
}
// lengths that get pruned instantiations (the large grids, where the pencil
// passes' arithmetic counts)
__host__ __device__ constexpr bool pz_len(int n) { return n >= 256; }
__host__ __device__ constexpr bool pz_cut(int n, int cut) { return 8 * cut < 3 * n; }

// Non-temporal loads and stores per pass, as bit masks over the passes.
// Measured at 512^3: non-temporal stores of every pass's output are faster
// (2% per step in FP64 and FP32 work for passes A and D, 2% more for B and
// C, neutral for E), also where the next pass reads the data back from the
// memory-side cache; non-temporal loads are slower (C and E by 3-7%).
enum NtPass { NT_A = 1, NT_B = 2, NT_C = 4, NT_D = 8, NT_E = 16 };
#ifndef GT_NT_ST
#define GT_NT_ST (NT_A | NT_B | NT_C | NT_D | NT_E)
#endif
#ifndef GT_NT_LD
#define GT_NT_LD 0
#endif
template <int PASS, typename V>
__device__ __forceinline__ V ld_row_p(const V* row, unsigned i) {
  if constexpr ((GT_NT_LD & PASS) != 0) {
    using E = decltype(row->x);
    typedef E E2 __attribute__((ext_vector_type(2)));
    const E2 x = __builtin_nontemporal_load(reinterpret_cast<const E2*>(reinterpret_cast<const char*>(row) + i * (unsigned)sizeof(V)));
    return V{x.x, x.y};
  } else {
    return ld_row(row, i);
  }
}
template <int PASS, typename V>
__device__ __forceinline__ void st_row_p(V* row, unsigned i, V v) {
  if constexpr ((GT_NT_ST & PASS) != 0) {
    using E = decltype(v.x);
    typedef E E2 __attribute__((ext_vector_type(2)));
    const E2 x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<E2*>(reinterpret_cast<char*>(row) + i * (unsigned)sizeof(V)));
  } else {
    st_row(row, i, v);
  }
}

// Packed C2R input of one field at element n of the sequence: modes above
// kc are zero, the imaginary parts of X[0] and X[N/2] are dropped (the C2R of
// fft_inv, src/fft_inv.f90:12).  For FP64 rows the loads are unconditional
// (a lane above kc reads element 0 of its row and discards it), so that a
// thread issues all its loads before it waits for any: a load under a
// branch is waited for inside the branch.
template <typename W, int N>
__device__ __forceinline__ cx<W> c2r_in(const cx<W>* r0, const cx<W>* r1, int n, int kc) {
  const int mm = (n <= N / 2) ? n : N - n;
  const bool ok = mm <= kc;
  const bool edge = (mm == 0 || mm == N / 2);
  cx<W> X0 = {W(0), W(0)}, X1 = {W(0), W(0)};
  if constexpr (sizeof(W) == 8) {
    const unsigned at = ok ? mm : 0;
    X0 = ld_row(r0, at);
    X1 = ld_row(r1, at);
    X0 = {ok ? X0.x : W(0), (ok && !edge) ? X0.y : W(0)};
    X1 = {ok ? X1.x : W(0), (ok && !edge) ? X1.y : W(0)};
  } else {
    // FP32 rows: occupancy hides the latency; predicated loads fetch less
    if (ok) {
      X0 = ld_row(r0, (unsigned)mm);
      X1 = ld_row(r1, (unsigned)mm);
    }
    if (edge) {
      X0.y = W(0);
      X1.y = W(0);
    }
  }
  return herm_pack(X0, X1, n > N / 2);
}

// c2r_in at position n = t + C of a thread's slot, for a compile-time
// offset C and 0 <= t < P: when the slot lies wholly below N/2 or wholly
// above it, the half (and so the index, the conjugation and the edge test)
// is known at compile time.  Same values as c2r_in.
template <typename W, int N, int P, int C>
__device__ __forceinline__ cx<W> c2r_in_at(const cx<W>* r0, const cx<W>* r1, int t, int kc) {
  if constexpr (C + P - 1 < N / 2) {
    const int mm = t + C;
    const bool ok = mm <= kc;
    cx<W> X0 = {W(0), W(0)}, X1 = {W(0), W(0)};
    if constexpr (sizeof(W) == 8) {
      const unsigned at = ok ? mm : 0;
      X0 = ld_row_p<NT_C>(r0, at);
      X1 = ld_row_p<NT_C>(r1, at);
      X0 = {ok ? X0.x : W(0), ok ? X0.y : W(0)};
      X1 = {ok ? X1.x : W(0), ok ? X1.y : W(0)};
    } else if (ok) {
      X0 = ld_row_p<NT_C>(r0, (unsigned)mm);
      X1 = ld_row_p<NT_C>(r1, (unsigned)mm);
    }
    if constexpr (C == 0) {
      if (t == 0) X0.y = X1.y = W(0);  // X[0]: imaginary parts dropped
    }
    return herm_pack(X0, X1, false);
  } else if constexpr (C > N / 2) {
    const int mm = N - C - t;  // in (0, N/2)
    const bool ok = mm <= kc;
    cx<W> X0 = {W(0), W(0)}, X1 = {W(0), W(0)};
    if constexpr (sizeof(W) == 8) {
      const unsigned at = ok ? mm : 0;
      X0 = ld_row_p<NT_C>(r0, at);
      X1 = ld_row_p<NT_C>(r1, at);
      X0 = {ok ? X0.x : W(0), ok ? X0.y : W(0)};
      X1 = {ok ? X1.x : W(0), ok ? X1.y : W(0)};
    } else if (ok) {
      X0 = ld_row_p<NT_C>(r0, (unsigned)mm);
      X1 = ld_row_p<NT_C>(r1, (unsigned)mm);
    }
    return herm_pack(X0, X1, true);
  } else {
    return c2r_in<W, N>(r0, r1, t + C, kc);
  }
}

}  // namespace gt
