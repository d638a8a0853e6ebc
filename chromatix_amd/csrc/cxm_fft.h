// Complex64 Stockham FFT building blocks for gfx950.
//
// One transform of length N is computed by T = N / E cooperating threads;
// thread t owns E values.  Layout invariant, before and after every
// transform:   v[k]  <->  X[t + k * T],  k = 0 .. E-1.
// Loads and stores against global memory in this layout are coalesced along
// t (consecutive lanes touch consecutive elements), for rows and for column
// strips alike.  Between radix stages values are exchanged through LDS; the
// last stage writes straight back to registers in the invariant layout, so a
// transform of s stages costs s-1 LDS round trips.
//
// Stage with radix R after a product Ns of earlier radices (Stockham
// auto-sort, decimation in time; Govindaraju et al., SC'08): thread butterfly
// m handles j = t + m*T; inputs are X[j + r*N/R], twiddled by
// W_{Ns*R}^{(j mod Ns) * r}; outputs go to (j / Ns) * Ns * R + j mod Ns + r*Ns.
// Each radix must divide E so that every thread owns whole butterflies.
//
// Twiddles come from the per-length FP64-built table of cxm_common.h,
// rounded once to f32.
#pragma once
#include "cxm_common.h"

namespace cxm {

// ---- compile-time trigonometry (for butterfly-internal constants) --------
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double cx_sin_series(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 30; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_reduce(double x) {  // to [-pi, pi]
  while (x > kPi) x -= 2.0 * kPi;
  while (x < -kPi) x += 2.0 * kPi;
  return x;
}
constexpr double cx_sin(double x) { return cx_sin_series(cx_reduce(x)); }
constexpr double cx_cos(double x) { return cx_sin(x + kPi / 2.0); }

// a * W_R^M with W_R = exp(DIR * 2 pi i / R); trivial angles handled exactly.
template <int R, int M, int DIR>
CXM_DEV float2 twc(float2 a) {
  constexpr int m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == R) {  // exp(DIR i pi/2)
    return mul_i<DIR>(a);
  } else if constexpr (2 * m == R) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * m == 3 * R) {
    return mul_i<-DIR>(a);
  } else {
    constexpr float c = (float)cx_cos(2.0 * kPi * m / R);
    constexpr float s = (float)(DIR * cx_sin(2.0 * kPi * m / R));
    return make_float2(__builtin_fmaf(a.x, c, -(a.y * s)), __builtin_fmaf(a.x, s, a.y * c));
  }
}

// ---- small DFTs, in place, natural order in and out ----------------------
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<2, DIR> {
  static CXM_DEV void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int DIR> struct Dft<4, DIR> {
  static CXM_DEV void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = mul_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int DIR> struct Dft<3, DIR> {
  static CXM_DEV void run(float2* a) {
    constexpr float s = (float)(0.86602540378443864676);  // sin(2pi/3)
    float2 b = cadd(a[1], a[2]);
    float2 d = csub(a[1], a[2]);
    float2 m = make_float2(__builtin_fmaf(-0.5f, b.x, a[0].x), __builtin_fmaf(-0.5f, b.y, a[0].y));
    // DIR * i * s * d
    float2 r = mul_i<DIR>(cscale(d, s));
    a[0] = cadd(a[0], b);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int DIR> struct Dft<5, DIR> {
  static CXM_DEV void run(float2* a) {
    constexpr float c1 = (float)0.30901699437494742410;   // cos(2pi/5)
    constexpr float c2 = (float)-0.80901699437494742410;  // cos(4pi/5)
    constexpr float s1 = (float)0.95105651629515357212;   // sin(2pi/5)
    constexpr float s2 = (float)0.58778525229247312917;   // sin(4pi/5)
    float2 b1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    float2 b2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    float2 x0 = a[0];
    const float2 p1 = make_float2(__builtin_fmaf(c2, b2.x, __builtin_fmaf(c1, b1.x, x0.x)),
                                  __builtin_fmaf(c2, b2.y, __builtin_fmaf(c1, b1.y, x0.y)));
    const float2 p2 = make_float2(__builtin_fmaf(c1, b2.x, __builtin_fmaf(c2, b1.x, x0.x)),
                                  __builtin_fmaf(c1, b2.y, __builtin_fmaf(c2, b1.y, x0.y)));
    float2 q1 = make_float2(__builtin_fmaf(s2, d2.x, s1 * d1.x), __builtin_fmaf(s2, d2.y, s1 * d1.y));
    float2 q2 = make_float2(__builtin_fmaf(-s1, d2.x, s2 * d1.x), __builtin_fmaf(-s1, d2.y, s2 * d1.y));
    q1 = mul_i<DIR>(q1);
    q2 = mul_i<DIR>(q2);
    a[0] = cadd(x0, cadd(b1, b2));
    a[1] = cadd(p1, q1);
    a[4] = csub(p1, q1);
    a[2] = cadd(p2, q2);
    a[3] = csub(p2, q2);
  }
};

// Composite R = P * Q:  n = n1 + P n2,  k = k1 + Q k2.
template <int P, int Q, int DIR>
struct DftPQ {
  static CXM_DEV void run(float2* a) {
    constexpr int R = P * Q;
    float2 b[P][Q];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) b[n1][n2] = a[n1 + P * n2];
      Dft<Q, DIR>::run(b[n1]);
    }
    twiddle_all<0, 0>(b);
    float2 c[P];
#pragma unroll
    for (int k1 = 0; k1 < Q; ++k1) {
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) c[n1] = b[n1][k1];
      Dft<P, DIR>::run(c);
#pragma unroll
      for (int k2 = 0; k2 < P; ++k2) a[k1 + Q * k2] = c[k2];
    }
    (void)R;
  }
  template <int N1, int K1>
  static CXM_DEV void twiddle_all(float2 (&b)[P][Q]) {
    if constexpr (N1 < P) {
      if constexpr (K1 < Q) {
        b[N1][K1] = twc<P * Q, N1 * K1, DIR>(b[N1][K1]);
        twiddle_all<N1, K1 + 1>(b);
      } else {
        twiddle_all<N1 + 1, 0>(b);
      }
    }
  }
};
template <int DIR> struct Dft<8, DIR> : DftPQ<4, 2, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftPQ<4, 4, DIR> {};

// ---- radix schedule --------------------------------------------------------
// Odd radices first, then the largest power-of-two radix dividing both the
// remaining length and E.
constexpr int pick_radix(int rem, int E) {
  if (rem % 5 == 0 && E % 5 == 0) return 5;
  if (rem % 3 == 0 && E % 3 == 0) return 3;
  if (rem % 16 == 0 && E % 16 == 0) return 16;
  if (rem % 8 == 0 && E % 8 == 0) return 8;
  if (rem % 4 == 0 && E % 4 == 0) return 4;
  if (rem % 2 == 0 && E % 2 == 0) return 2;
  return -1;
}
constexpr bool schedule_ok(int N, int E) {
  int ns = 1;
  while (ns < N) {
    int r = pick_radix(N / ns, E);
    if (r <= 0) return false;
    ns *= r;
  }
  return ns == N;
}
constexpr int num_stages(int N, int E) {
  int ns = 1, s = 0;
  while (ns < N) { ns *= pick_radix(N / ns, E); ++s; }
  return s;
}

// LDS slot of element e of a transform of length N: a stride S between
// elements (S column transforms interleaved, slot = row * S + column) and a
// one-row pad every 16 rows, which breaks the power-of-two bank aliasing of
// the radix-16 scatter (without it the S lanes of a row and the 16 rows a
// wave writes land on the same banks: 2.75x the ideal LDS cycles for
// 1024-point columns of 4).  Lengths 5*2^a and 3*2^a with S > 1 keep the
// plain stride, which conflicts less for their radix-5/3 stages.  (An XOR
// row swizzle removes the conflicts too, but its address arithmetic costs
// more registers than this pad.)
template <int N, int S>
constexpr bool lds_padded() { return S == 1 || (N & (N - 1)) == 0; }
template <int N, int S>
CXM_DEV int lds_at(int e) {
  if constexpr (lds_padded<N, S>()) return (e + (e >> 4)) * S;
  else return e * S;
}
// lds_at(a + b) where (a mod 16) + (b mod 16) < 16, so that the row pad
// count of a + b is that of a plus that of b: the slot splits into a runtime
// part and a compile-time part that becomes the instruction's immediate
// offset (one address register per butterfly instead of one per value).
// Callers: a = t < T, b = e * T (T a power of two); a = q * Ns * R + k with
// k < Ns, b = r * Ns (Ns, R powers of two).
template <int N, int S>
CXM_DEV int lds_at_split(int a, int b) {
  if constexpr (lds_padded<N, S>() && (N & (N - 1)) == 0) return lds_at<N, S>(a) + lds_at<N, S>(b);
  else return lds_at<N, S>(a + b);
}
template <int N, int S>
constexpr int lds_span() {  // float2 slots occupied by S interleaved transforms
  if constexpr (lds_padded<N, S>()) return (N + N / 16) * S;
  else return N * S;
}

// One Stockham stage.  `lds` points at this transform's element 0.
// SYNC: 0 -> wave-local transform (T <= 64 and one transform per wave group),
// 1 -> block-wide barrier.
// Twiddles of the stage that starts after a product Ns of earlier radices:
// butterfly m, input r >= 1 is multiplied by W_{Ns*R}^{(j mod Ns) * r}.  A
// stage loads the next stage's twiddles before its LDS exchange, so that the
// loads are in flight during the barrier rather than exposed after it.
template <int N, int E, int Ns>
struct StageTw {
  static constexpr bool active = Ns > 1 && Ns < N;
  static constexpr int R = (Ns < N) ? pick_radix(N / Ns, E) : 1;
  static constexpr int COUNT = active ? (E / R) * (R - 1) : 1;
  float2 w[COUNT];
};

template <int N, int E, int DIR, int Ns>
CXM_DEV void load_stage_tw(StageTw<N, E, Ns>& st, int t, const float2* __restrict__ tw) {
  if constexpr (StageTw<N, E, Ns>::active) {
    constexpr int R = StageTw<N, E, Ns>::R;
    constexpr int M = E / R;
    constexpr int T = N / E;
    constexpr int twstep = N / (Ns * R);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int k = (t + m * T) % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 x = tw[k * r * twstep];
        if constexpr (DIR > 0) x.y = -x.y;
        st.w[m * (R - 1) + r - 1] = x;
      }
    }
  }
}

// Butterflies of one stage (twiddles applied), results to LDS unless it is
// the last stage, which leaves them in registers in the invariant layout.
template <int N, int E, int DIR, int Ns, int R, int S>
CXM_DEV void stage_body(float2 (&v)[E], float2* lds, int t, const StageTw<N, E, Ns>& w) {
  constexpr int T = N / E;
  constexpr int M = E / R;
  constexpr bool last = (Ns * R == N);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = t + m * T;
    const int k = j % Ns;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[m + r * M];
    if constexpr (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w.w[m * (R - 1) + r - 1]);
    }
    if constexpr (R == 2) Dft<2, DIR>::run(a);
    else if constexpr (R == 3) Dft<3, DIR>::run(a);
    else if constexpr (R == 4) Dft<4, DIR>::run(a);
    else if constexpr (R == 5) Dft<5, DIR>::run(a);
    else if constexpr (R == 8) Dft<8, DIR>::run(a);
    else Dft<16, DIR>::run(a);
    if constexpr (last) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m + r * M] = a[r];
    } else {
      // base = q * Ns * R + k with k < Ns: base + r * Ns never carries past
      // a 16-row boundary that r * Ns does not (powers of two)
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lds_at_split<N, S>(base, r * Ns)] = a[r];
    }
  }
}

// Sync, read back in the invariant layout, sync.
template <int N, int E, int S, bool SYNC>
CXM_DEV void stage_exchange(float2 (&v)[E], float2* lds, int t) {
  constexpr int T = N / E;
  if constexpr (SYNC) block_sync();
  else wave_sync();
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = lds[lds_at_split<N, S>(t, e * T)];
  if constexpr (SYNC) block_sync();
  else wave_sync();
}

template <int N, int E, int DIR, int Ns, int S, bool SYNC>
CXM_DEV void stockham_from(float2 (&v)[E], float2* lds, int t, const float2* __restrict__ tw,
                           const StageTw<N, E, Ns>& w) {
  if constexpr (Ns < N) {
    constexpr int R = pick_radix(N / Ns, E);
    static_assert(R > 0, "FFT length / thread split has no radix schedule");
    stage_body<N, E, DIR, Ns, R, S>(v, lds, t, w);
    if constexpr (Ns * R < N) {
      StageTw<N, E, Ns * R> wn;
      load_stage_tw<N, E, DIR, Ns * R>(wn, t, tw);
      stage_exchange<N, E, S, SYNC>(v, lds, t);
      stockham_from<N, E, DIR, Ns * R, S, SYNC>(v, lds, t, tw, wn);
    }
  }
}

// The twiddles of every stage from Ns on, held in registers (for loops that
// transform many rows with one thread layout).
template <int N, int E, int Ns, bool END = (Ns >= N)>
struct AllTw;
template <int N, int E, int Ns>
struct AllTw<N, E, Ns, true> {};
template <int N, int E, int Ns>
struct AllTw<N, E, Ns, false> {
  static constexpr int R = pick_radix(N / Ns, E);
  StageTw<N, E, Ns> here;
  AllTw<N, E, Ns * R> rest;
};

template <int N, int E, int DIR, int Ns>
CXM_DEV void load_all_tw(AllTw<N, E, Ns>& a, int t, const float2* __restrict__ tw) {
  if constexpr (Ns < N) {
    load_stage_tw<N, E, DIR, Ns>(a.here, t, tw);
    load_all_tw<N, E, DIR, Ns * AllTw<N, E, Ns>::R>(a.rest, t, tw);
  }
}

template <int N, int E, int DIR, int Ns, int S, bool SYNC>
CXM_DEV void stockham_pre(float2 (&v)[E], float2* lds, int t, const AllTw<N, E, Ns>& all) {
  if constexpr (Ns < N) {
    constexpr int R = AllTw<N, E, Ns>::R;
    stage_body<N, E, DIR, Ns, R, S>(v, lds, t, all.here);
    if constexpr (Ns * R < N) {
      stage_exchange<N, E, S, SYNC>(v, lds, t);
      stockham_pre<N, E, DIR, Ns * R, S, SYNC>(v, lds, t, all.rest);
    }
  }
}

// Full transform with twiddles loaded beforehand by load_all_tw<N, E, DIR, 1>.
template <int N, int E, int DIR, int S, bool SYNC>
CXM_DEV void fft_pre(float2 (&v)[E], float2* lds, int t, const AllTw<N, E, 1>& all) {
  static_assert(N % E == 0, "E must divide N");
  static_assert(schedule_ok(N, E), "no radix schedule for (N, E)");
  stockham_pre<N, E, DIR, 1, S, SYNC>(v, lds, t, all);
}

// Full transform, unnormalised.  DIR = -1: forward (exp(-2 pi i nk/N)),
// DIR = +1: inverse without the 1/N.
template <int N, int E, int DIR, int S, bool SYNC>
CXM_DEV void fft(float2 (&v)[E], float2* lds, int t, const float2* __restrict__ tw) {
  static_assert(N % E == 0, "E must divide N");
  static_assert(schedule_ok(N, E), "no radix schedule for (N, E)");
  const StageTw<N, E, 1> none{};
  stockham_from<N, E, DIR, 1, S, SYNC>(v, lds, t, tw, none);
}

}  // namespace cxm
