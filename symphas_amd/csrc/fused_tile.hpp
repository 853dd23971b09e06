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

#include <type_traits>

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

// Keep v in a VGPR from here on (an empty asm that "modifies" it), so the
// compiler does not re-copy a scalar into a VGPR at every use.
template <typename T>
__device__ __forceinline__ void pin_vgpr(T& v) {
  asm volatile("" : "+v"(v));
}

// Calls f(std::integral_constant<int, U>) for U = B .. E-1, so that ring
// slots and stream phases derived from U are compile-time constants.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Box mode of the stream (a regional grid's box, X x Y with a layer of
// `depth` sites): rows outside [0, Y) and columns outside [0, X) read 0 (no
// periodic wrap), and after each stage a site outside the interior
// [depth, X-depth) x [depth, Y-depth) keeps its input value, so the layer
// stays frozen.  cols: bit j set when the lane's value j lies in an interior
// column.
struct BoxMask {
  int depth = 0;
  unsigned cols = 0;
};

// Process one strip x chunk.  col: first column this lane loads (wrapped into
// [0, X); in box mode it may lie outside); y0/ty: first output row and row
// count of the chunk.
//
// The chunk streams NI = ty + 2*K*R rows: stage s has valid input from
// stream row 2*R*s on, and output row o leaves the last stage with stream
// row o + 2*K*R.  The schedule is split so that the hot loop carries no
// per-row control flow:
//   - fill (rows 0 .. 2*K*R-1, unrolled): row i runs the stages whose input
//     is valid (s <= i / (2*R)) and stores nothing;
//   - steady (groups of D rows): every stage, every row stored;
//   - last group: the same with a uniform check per row.
// Rows are prefetched D ahead into a register ring (the row in slot
// i mod D is replaced by row i + D as soon as it is read).  Loads wrap
// modulo Y like the rows themselves, so every address is inside the field.
template <typename T, int VEC, int K, int R, bool HASA, int POLY, bool BOX = false>
__device__ __forceinline__ void stream_tile(const T* __restrict__ in,
                                            T* __restrict__ out, int X, int Y,
                                            int y0, int ty, int col,
                                            bool store_lane,
                                            const FusedCoef<T>& c,
                                            const BoxMask& bm = BoxMask{}) {
  constexpr int LAG = K * R;
  constexpr int FILL = 2 * LAG;
  constexpr int D = 9;  // a multiple of 3: slot and phase repeat together
  const int NI = ty + FILL;
  const int ybase = y0 - LAG;  // row of stream row 0
  int rload = BOX ? ybase : ybase % Y;
  if (!BOX && rload < 0) rload += Y;

  Stage<T, VEC, R, HASA, POLY> st[K];
#pragma unroll
  for (int s = 0; s < K; ++s) st[s].clear();

  // A VOP3 reads one scalar operand: the coefficient that meets another
  // scalar in the first FMA of a Horner chain is held in a VGPR.
  FusedCoef<T> cv = c;
  if constexpr (R == 2) pin_vgpr(cv.p[POLY - 2]);
  if constexpr (HASA) pin_vgpr(cv.q[POLY - 1]);

  // rows are addressed by a wave-uniform descriptor per row and the lane's
  // byte offset in it; lanes that do not store get an offset past the row
  const int row_bytes = X * (int)sizeof(T);
  const bool col_in = !BOX || (col >= 0 && col < X);
  const unsigned ld_off = col_in ? (unsigned)col * (unsigned)sizeof(T) : 0x80000000u;
  const unsigned st_off = store_lane ? ld_off : 0x80000000u;
  T ring[D][VEC];
  auto load_row = [&](T (&b)[VEC]) {
    if constexpr (BOX) {
      // a row outside the box: an empty range, every lane reads 0
      const bool row_in = (unsigned)rload < (unsigned)Y;
      buf_load<T, VEC>(buffer_rsrc(in + (int64_t)(row_in ? rload : 0) * X,
                                   row_in ? row_bytes : 0),
                       ld_off, 0, b);
      ++rload;
    } else {
      buf_load<T, VEC>(buffer_rsrc(in + (int64_t)rload * X, row_bytes), ld_off, 0, b);
      rload = rload + 1 == Y ? 0 : rload + 1;
    }
  };
  static_for<0, D>([&](auto u) { load_row(ring[decltype(u)::value]); });

  // stream row `row` (compile-time slot/phase index I) through stages
  // 0..NS-1; in box mode a stage's output outside the interior keeps the
  // stage's input value (P[S0] of a radius-two stage, P[S1] of radius one)
  auto run_row = [&](auto ic, auto nsc, int row, T (&v)[VEC]) {
    constexpr int I = decltype(ic)::value, NS = decltype(nsc)::value;
#pragma unroll
    for (int j = 0; j < VEC; ++j) v[j] = ring[I % D][j];
    static_for<0, NS>([&](auto sc) {
      constexpr int S = decltype(sc)::value;
      st[S].template push<I % 3>(cv, v);
      if constexpr (BOX) {
        const int yo = ybase + row - R * (S + 1);
        const bool row_in = yo >= bm.depth && yo < Y - bm.depth;
        constexpr int KEEP = R == 2 ? I % 3 : (I + 1) % 3;
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          if (!(row_in && ((bm.cols >> j) & 1u))) v[j] = st[S].P[KEEP][j];
      }
    });
  };

  static_for<0, FILL>([&](auto ic) {
    constexpr int I = decltype(ic)::value;
    constexpr int NS = I / (2 * R) + 1 < K ? I / (2 * R) + 1 : K;
    T v[VEC];
    run_row(ic, std::integral_constant<int, NS>{}, I, v);
    load_row(ring[I % D]);
  });

  T* dst = out + (int64_t)y0 * X;
  int i = FILL;
  // two groups per iteration while three fit, then one
  for (; i + 3 * D <= NI; i += 2 * D) {
    static_for<0, 2 * D>([&](auto uc) {
      constexpr int I = FILL + decltype(uc)::value;
      T v[VEC];
      run_row(std::integral_constant<int, I>{}, std::integral_constant<int, K>{},
              i + decltype(uc)::value, v);
      load_row(ring[I % D]);
      buf_store<T, VEC>(buffer_rsrc(dst, row_bytes), st_off, 0, v);
      dst += X;
    });
  }
  for (; i + 2 * D <= NI; i += D) {
    static_for<0, D>([&](auto uc) {
      constexpr int I = FILL + decltype(uc)::value;
      T v[VEC];
      run_row(std::integral_constant<int, I>{}, std::integral_constant<int, K>{},
              i + decltype(uc)::value, v);
      load_row(ring[I % D]);
      buf_store<T, VEC>(buffer_rsrc(dst, row_bytes), st_off, 0, v);
      dst += X;
    });
  }
  for (; i < NI; i += D) {
    static_for<0, D>([&](auto uc) {
      constexpr int U = decltype(uc)::value, I = FILL + U;
      if (i + U < NI) {
        T v[VEC];
        run_row(std::integral_constant<int, I>{}, std::integral_constant<int, K>{}, i + U, v);
        if (i + U + D < NI) load_row(ring[I % D]);
        buf_store<T, VEC>(buffer_rsrc(dst, row_bytes), st_off, 0, v);
        dst += X;
      }
    });
  }
}

}  // namespace sps
