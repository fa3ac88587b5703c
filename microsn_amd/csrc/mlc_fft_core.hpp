// mlc_fft_core.hpp -- radix butterflies and the Stockham index map of the
// library's FFTs.  Written once, compiled both for gfx950 (inside the LDS
// kernels of mlc_fft.hip) and for the host (a CPU harness drives the same
// functions to check the index arithmetic against a reference DFT).
//
// One Stockham stage of radix R over a length-L sequence, after stages whose
// radices multiply to Ns, maps butterfly j (0 <= j < L/R) as
//   inputs   x[j + r*L/R],           r = 0..R-1
//   twiddle  x_r *= W_L^(r*k*L/(Ns*R)),  k = j mod Ns,  W_L = exp(-+2 pi i/L)
//   outputs  y[(j/Ns)*Ns*R + k + q*Ns] = sum_r x_r W_R^(r*q),  q = 0..R-1
// (Stockham autosort, e.g. Van Loan, "Computational Frameworks for the FFT",
// sec. 1.7; the output is in natural order after the last stage).
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MLC_HD __host__ __device__ __forceinline__
#else
#define MLC_HD inline
#endif

namespace mlc {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <class T>
MLC_HD cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <class T>
MLC_HD cpx<T> add(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <class T>
MLC_HD cpx<T> sub(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <class T>
MLC_HD cpx<T> mul(cpx<T> a, cpx<T> b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <class T>
MLC_HD cpx<T> mulc(cpx<T> a, cpx<T> b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
template <class T>
MLC_HD cpx<T> conj(cpx<T> a) { return mk(a.x, -a.y); }
template <class T>
MLC_HD cpx<T> scale(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <class T, bool INV>
MLC_HD cpx<T> rot90(cpx<T> a) { return INV ? mk(-a.y, a.x) : mk(a.y, -a.x); }

// ------------------------------------------------------------------ DFTs
// In-place R-point DFT of v[0..R): forward uses W_R = exp(-2 pi i/R),
// inverse its conjugate.  Unnormalised.
template <class T, bool INV>
MLC_HD void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = add(a, b);
  v[1] = sub(a, b);
}

template <class T, bool INV>
MLC_HD void dft4(cpx<T>* v) {
  cpx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
  cpx<T> t2 = add(v[1], v[3]), t3 = rot90<T, INV>(sub(v[1], v[3]));
  v[0] = add(t0, t2);
  v[2] = sub(t0, t2);
  v[1] = add(t1, t3);
  v[3] = sub(t1, t3);
}

template <class T, bool INV>
MLC_HD void dft8(cpx<T>* v) {
  // radix-2 split: even and odd 4-point DFTs, odd part twiddled by W_8^q
  const T h = (T)0.70710678118654752440084436210484903928;
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, INV>(e);
  dft4<T, INV>(o);
  // W_8^1 = (1 -+ i)/sqrt2, W_8^2 = -+i, W_8^3 = (-1 -+ i)/sqrt2
  cpx<T> o1 = INV ? mk((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                  : mk((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
  cpx<T> o2 = rot90<T, INV>(o[2]);
  cpx<T> o3 = INV ? mk(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h)
                  : mk((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
  v[0] = add(e[0], o[0]);
  v[4] = sub(e[0], o[0]);
  v[1] = add(e[1], o1);
  v[5] = sub(e[1], o1);
  v[2] = add(e[2], o2);
  v[6] = sub(e[2], o2);
  v[3] = add(e[3], o3);
  v[7] = sub(e[3], o3);
}

template <class T, bool INV>
MLC_HD void dft3(cpx<T>* v) {
  // y0 = a+b+c; y1,2 = a - (b+c)/2 -+ i*(sqrt3/2)*(b-c) (forward)
  const T s = (T)0.86602540378443864676372317075293618347;
  cpx<T> t = add(v[1], v[2]);
  cpx<T> d = sub(v[1], v[2]);
  cpx<T> m = mk(v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y);
  cpx<T> r = rot90<T, INV>(scale(d, s));  // -i*s*d forward
  v[0] = add(v[0], t);
  v[1] = add(m, r);
  v[2] = sub(m, r);
}

template <class T, bool INV>
MLC_HD void dft5(cpx<T>* v) {
  // direct form with c1 = cos(2pi/5), c2 = cos(4pi/5), s1 = sin(2pi/5), s2 = sin(4pi/5)
  const T c1 = (T)0.30901699437494742410229341718281905886;
  const T c2 = (T)-0.80901699437494742410229341718281905886;
  const T s1 = (T)0.95105651629515357211643933337938214340;
  const T s2 = (T)0.58778525229247312916870595463907276860;
  cpx<T> a1 = add(v[1], v[4]), b1 = sub(v[1], v[4]);
  cpx<T> a2 = add(v[2], v[3]), b2 = sub(v[2], v[3]);
  cpx<T> m1 = mk(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  cpx<T> m2 = mk(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  // forward: y1 = m1 - i(s1 b1 + s2 b2), y4 = m1 + i(...); y2 = m2 - i(s2 b1 - s1 b2), y3 = m2 + i(...)
  cpx<T> n1 = rot90<T, INV>(mk(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  cpx<T> n2 = rot90<T, INV>(mk(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = add(v[0], add(a1, a2));
  v[1] = add(m1, n1);
  v[4] = sub(m1, n1);
  v[2] = add(m2, n2);
  v[3] = sub(m2, n2);
}

template <class T, bool INV>
MLC_HD void dft16(cpx<T>* v) {
  // 16 = 4 x 4: DFT4 over n2 for each n1 (x[n1 + 4 n2]), twiddle W_16^(n1 k2),
  // DFT4 over n1 for each k2; output X[k2 + 4 k1].
  const T c1 = (T)0.92387953251128675612818318939678828682;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977172845998403039886676;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210484903928;
  cpx<T> y[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    cpx<T> a[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4<T, INV>(a);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) y[n1][k2] = a[k2];
  }
  // W_16^m = cos(2 pi m/16) -+ i sin(2 pi m/16), m = n1*k2 in {1,2,3,4,6,9}
  const T cs[10] = {1, c1, h, s1, 0, 0, -h, 0, 0, -c1};
  const T sn[10] = {0, s1, h, c1, 1, 0, h, 0, 0, -s1};
#pragma unroll
  for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
    for (int k2 = 1; k2 < 4; ++k2) {
      const T c = cs[n1 * k2], sv = INV ? -sn[n1 * k2] : sn[n1 * k2];
      const cpx<T> a = y[n1][k2];
      y[n1][k2] = mk(a.x * c + a.y * sv, a.y * c - a.x * sv);  // a * (c - i sv)
    }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    cpx<T> a[4] = {y[0][k2], y[1][k2], y[2][k2], y[3][k2]};
    dft4<T, INV>(a);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = a[k1];
  }
}

template <class T, int R, bool INV>
MLC_HD void dft(cpx<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else if constexpr (R == 4) dft4<T, INV>(v);
  else if constexpr (R == 5) dft5<T, INV>(v);
  else if constexpr (R == 8) dft8<T, INV>(v);
  else if constexpr (R == 16) dft16<T, INV>(v);
}

// ------------------------------------------------------------------ stage
// Index map of the sequence in memory (identity; the LDS kernels use it).
struct NoPad {
  MLC_HD int operator()(int i) const { return i; }
};
// Twiddle source W_L^m = exp(-2 pi i m/L), m in [0, L): a full table.
template <class T>
struct TableTw {
  const cpx<T>* tw;
  MLC_HD cpx<T> operator()(int m) const { return tw[m]; }
};
// The same from a quarter-wave table q[r] = W_L^r, r in [0, L/4] (L % 4 == 0):
// W_L^(a L/4 + r) = (-i)^a W_L^r.
template <class T>
struct QuarterTw {
  const cpx<T>* q;
  int QL;  // L/4
  MLC_HD cpx<T> operator()(int m) const {
    const int a = (m >= QL) + (m >= 2 * QL) + (m >= 3 * QL);
    const cpx<T> w = q[m - a * QL];
    switch (a) {
      case 0: return w;
      case 1: return mk(w.y, -w.x);
      case 2: return mk(-w.x, -w.y);
      default: return mk(-w.y, w.x);
    }
  }
};

// Loads, twiddles and transforms butterfly j into v[0..R) (conjugate
// twiddles for INV).
template <class T, int R, bool INV, class TW, class PAD = NoPad>
MLC_HD void stage_load(const cpx<T>* x, int L, int Ns, int j, const TW& tw, cpx<T>* v,
                       PAD pad = PAD()) {
  const int stride = L / R;
  const int k = j % Ns;
  const int tstep = k * (L / (Ns * R));
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = x[pad(j + r * stride)];
  if (k != 0) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cpx<T> w = tw(r * tstep);
      v[r] = INV ? mulc(v[r], w) : mul(v[r], w);
    }
  }
  dft<T, R, INV>(v);
}

template <class T, int R, class PAD = NoPad>
MLC_HD void stage_store(cpx<T>* y, int Ns, int j, const cpx<T>* v, PAD pad = PAD()) {
  const int k = j % Ns;
  const int base = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int q = 0; q < R; ++q) y[pad(base + q * Ns)] = v[q];
}

// ------------------------------------------------------------------ real FFT
// A real sequence x of even length N = 2M is transformed through the complex
// length-M sequence z[n] = x[2n] + i x[2n+1] with Z = DFT_M(z).  With
// E = (Z[k] + conj(Z[M-k]))/2, O = -i (Z[k] - conj(Z[M-k]))/2 and
// w = W_N^k = exp(-2 pi i k/N):
//   X[k] = E + w O,   X[M-k] = conj(E - w O)          (k = 0..M/2, Z[M] = Z[0])
template <class T>
MLC_HD void r2c_post(cpx<T> Zk, cpx<T> Zm, cpx<T> w, cpx<T>* Xk, cpx<T>* Xm) {
  cpx<T> zc = conj(Zm);
  cpx<T> E = scale(add(Zk, zc), (T)0.5);
  cpx<T> D = sub(Zk, zc);
  cpx<T> O = mk(D.y * (T)0.5, -D.x * (T)0.5);  // -i D / 2
  cpx<T> wO = mul(w, O);
  *Xk = add(E, wO);
  *Xm = conj(sub(E, wO));
}

// Inverse: from the half spectrum X[0..M] of a real sequence, the complex
// input of the unnormalised inverse DFT_M whose output is
// z[n] = x[2n] + i x[2n+1] (x unnormalised: x[n] = sum_k X[k] e^{+2 pi i kn/N}):
//   Zin[k]   = (X[k] + conj(X[M-k])) + i (X[k] - conj(X[M-k])) conj(w)
//   Zin[M-k] = conj((X[k] + conj(X[M-k])) - i (X[k] - conj(X[M-k])) conj(w))
template <class T>
MLC_HD void c2r_pre(cpx<T> Xk, cpx<T> Xm, cpx<T> w, cpx<T>* Zk, cpx<T>* Zm) {
  cpx<T> xc = conj(Xm);
  cpx<T> S = add(Xk, xc);
  cpx<T> D = mulc(sub(Xk, xc), w);  // (X[k]-conj(X[M-k])) * conj(w)
  cpx<T> iD = mk(-D.y, D.x);
  *Zk = add(S, iD);
  *Zm = conj(sub(S, iD));
}

// Radix plan: factors of L, largest first among {16, 8, 4, 2} (none above
// max_radix) then 5, 3.
struct Plan {
  int L = 0;
  int n = 0;
  int r[24] = {0};
};

inline bool make_plan(int L, Plan* p, int max_radix = 16) {
  p->L = L;
  p->n = 0;
  int m = L;
  while (max_radix >= 16 && m % 16 == 0) { p->r[p->n++] = 16; m /= 16; }
  while (m % 8 == 0) { p->r[p->n++] = 8; m /= 8; }
  while (m % 4 == 0) { p->r[p->n++] = 4; m /= 4; }
  while (m % 2 == 0) { p->r[p->n++] = 2; m /= 2; }
  while (m % 5 == 0) { p->r[p->n++] = 5; m /= 5; }
  while (m % 3 == 0) { p->r[p->n++] = 3; m /= 3; }
  return m == 1 && p->n <= 24;
}

}  // namespace mlc
