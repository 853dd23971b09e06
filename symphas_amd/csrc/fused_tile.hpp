// fused_tile.hpp - the row-streaming, temporally blocked 2D step on one tile.
//
// A wave owns a strip of 64*VEC columns (VEC consecutive values per lane) and
// streams the rows of one chunk of the field through K pipelined time steps.
// Neighbours along x come from adjacent lanes through DPP; neighbours along y
// are the previous/next rows kept in registers, rotated by unrolling the row
// loop by three.  Each stage applies one step of the canonical form in the
// folded coefficients of FusedCoef (S = sum of the 4 neighbours):
//   radius 2:  mu~  = S + psi*p(psi)          (ks*S when p has degree > 2)
//              psi' = psi + q(psi) + k2*(S[mu~] - 4*mu~)
//   radius 1:  psi' = psi + q(psi) + k1*(S - 4*psi)
// so stage s lags stage s-1 by R rows and loses R columns at each strip edge.
// HL lanes at each side of a strip only feed the neighbours; the strip's
// 64-2*HL inner lanes store.  Any X that is a multiple of VEC.
#pragma once

#include "common.hpp"
#include "lanes.hpp"

namespace sps {

template <typename T, int VEC, int R, bool HASA, int POLY>
struct Stage {
  T P[3][VEC];  // input rows r-2, r-1, r (slots rotate with the phase)
  T M[3][VEC];  // mu rows r-3, r-2, r-1 (radius 2 only)

  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        P[s][j] = T(0);
        M[s][j] = T(0);
      }
  }

  // Sum of the 4 neighbours of the row held in slot C, with rows U and D
  // above and below; each lane move pairs with a register operand.
  template <int C, int U, int D>
  __device__ __forceinline__ void nsum(const T (&A)[3][VEC],
                                       T (&s)[VEC]) const {
    const T left0 = from_prev_lane(A[C][VEC - 1]);
    const T rightN = from_next_lane(A[C][0]);
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const T l = j == 0 ? left0 : A[C][j - 1];
      const T r = j == VEC - 1 ? rightN : A[C][j + 1];
      s[j] = (l + r) + (A[U][j] + A[D][j]);
    }
  }

  // Push input row (stream phase PH); v becomes this stage's output row.
  template <int PH>
  __device__ __forceinline__ void push(const FusedCoef<T>& c, T (&v)[VEC]) {
    constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;
#pragma unroll
    for (int j = 0; j < VEC; ++j) P[S2][j] = v[j];
    T s[VEC];
    nsum<S1, S2, S0>(P, s);
    if constexpr (R == 1) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const T x = P[S1][j];
        T base = x;
        if constexpr (HASA) base = base + horner<POLY>(c.q, x);
        v[j] = fmad(c.k1, fmad(T(-4), x, s[j]), base);
      }
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const T x = P[S1][j];
        const T t = horner<POLY - 1>(c.p, x);
        M[S2][j] = fmad(t, x, POLY == 3 ? s[j] : c.ks * s[j]);
      }
      nsum<S1, S2, S0>(M, s);
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        T base = P[S0][j];
        if constexpr (HASA) base = base + horner<POLY>(c.q, base);
        v[j] = fmad(c.k2, fmad(T(-4), M[S1][j], s[j]), base);
      }
    }
  }
};

// Process one strip x chunk.  col: first column this lane loads (already
// wrapped into [0, X)); y0/ty: first output row and row count of the chunk.
template <typename T, int VEC, int K, int R, bool HASA, int POLY>
__device__ __forceinline__ void stream_tile(const T* __restrict__ in,
                                            T* __restrict__ out, int X, int Y,
                                            int y0, int ty, int col,
                                            bool store_lane,
                                            const FusedCoef<T>& c) {
  constexpr int LAG = K * R;
  const int NI = ty + 2 * LAG;
  int rload = (y0 - LAG) % Y;
  if (rload < 0) rload += Y;

  Stage<T, VEC, R, HASA, POLY> st[K];
#pragma unroll
  for (int s = 0; s < K; ++s) st[s].clear();

  T bufA[3][VEC], bufB[3][VEC];

  auto load3 = [&](T (&buf)[3][VEC], int i) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (i + q < NI) {
        load_vec<VEC>(in + (int64_t)rload * X + col, buf[q]);
        rload = rload + 1 == Y ? 0 : rload + 1;
      }
    }
  };
  auto process3 = [&](T (&buf)[3][VEC], int i) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (i + q < NI) {
        T v[VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) v[j] = buf[q][j];
#pragma unroll
        for (int s = 0; s < K; ++s) {
          // stage s has valid input from stream row 2*R*s on (each stage
          // delays by R rows and needs 2*R rows of fill); before that it
          // would only turn garbage into garbage, so it is skipped.  Its
          // valid outputs depend only on pushes of valid rows.
          if (s > 0 && i + q < 2 * R * s) continue;
          if (q == 0) st[s].template push<0>(c, v);
          if (q == 1) st[s].template push<1>(c, v);
          if (q == 2) st[s].template push<2>(c, v);
        }
        const int orow = i + q - 2 * LAG;
        if (orow >= 0 && store_lane)
          store_vec<VEC>(out + (int64_t)(y0 + orow) * X + col, v);
      }
    }
  };

  load3(bufA, 0);
  for (int i = 0; i < NI; i += 6) {
    load3(bufB, i + 3);
    process3(bufA, i);
    load3(bufA, i + 6);
    process3(bufB, i + 3);
  }
}

}  // namespace sps
