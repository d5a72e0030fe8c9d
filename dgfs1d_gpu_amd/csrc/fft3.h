// Register-resident mixed-radix DFTs and the LDS-staged "partial 3D FFT"
// engine shared by every collision kernel.
//
// RegDFT<T,L,DIR>: out[k] = sum_j in[S*j+O] * exp(DIR*2*pi*i*j*k/L), natural
// order, fully unrolled (radix 4/2/3/5/7 decimation in time).  All twiddles
// are compile-time literals (cx_unit), trivial angles are special-cased.
//
// Fft3Engine<T,N,S,NT,DIR>: one work-group transforms an N^3 complex volume
// X[a1][a2][a3] along a3, a2, a1 (in that order) but keeps only the outputs
// whose a3-frequency/point o3 satisfies o3 = S*j + r (residue r = per-WG),
// which costs 1/S of a full transform (decimation in frequency on the first
// axis).  Pass 1 streams pencils from global memory through a caller load
// functor, pass 2 and pass 3 work in LDS (N*N*(N/S+1) complex), and pass 3
// hands each thread one finished a1-pencil (o1 = 0..N-1) in registers.
#pragma once

#include "dgfs_common.h"

namespace dgfs {

// ------------------------------------------------------------ twiddle multiply
// a * exp(2*pi*i*E/L) with E, L compile-time.
template <typename T, long E, long L>
DGFS_DEV cplx<T> twiddle(cplx<T> a) {
    constexpr long e = ((E % L) + L) % L;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (2 * e == L) {
        return mk(-a.x, -a.y);
    } else if constexpr (4 * e == L) {
        return mul_i<1>(a);
    } else if constexpr (4 * e == 3 * L) {
        return mul_i<-1>(a);
    } else if constexpr ((8 * e) % L == 0) {
        // odd multiple of pi/4: (+-1 +- i)/sqrt(2)
        constexpr cs_pair w = cx_unit(e, L);
        constexpr T h = T(0.70710678118654752440084436210485);
        constexpr T sc = w.c > 0 ? h : -h;
        constexpr T ss = w.s > 0 ? h : -h;
        // (x + iy)(sc + i ss)
        if constexpr (sc == ss) {
            return mk(sc * (a.x - a.y), sc * (a.x + a.y));
        } else {
            return mk(sc * (a.x + a.y), sc * (a.y - a.x));
        }
    } else {
        constexpr cs_pair w = cx_unit(e, L);
        return a * mk(T(w.c), T(w.s));
    }
}

// --------------------------------------------------------------- butterflies
template <typename T, int R, int DIR>
DGFS_DEV void butterfly(cplx<T>* t) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        cplx<T> a = t[0], b = t[1];
        t[0] = a + b;
        t[1] = a - b;
    } else if constexpr (R == 4) {
        cplx<T> a = t[0] + t[2], b = t[0] - t[2];
        cplx<T> c = t[1] + t[3], d = mul_i<DIR>(t[1] - t[3]);
        t[0] = a + c;
        t[2] = a - c;
        t[1] = b + d;
        t[3] = b - d;
    } else if constexpr (R == 3) {
        constexpr T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
        cplx<T> s = t[1] + t[2], d = t[1] - t[2];
        cplx<T> m = t[0] - T(0.5) * s;
        cplx<T> e = mul_i<DIR>(h * d);
        t[0] = t[0] + s;
        t[1] = m + e;
        t[2] = m - e;
    } else {
        // direct R-point DFT (R = 5, 7, other primes)
        cplx<T> in[R];
        sfor<0, R>([&](auto I) { in[I] = t[I]; });
        sfor<0, R>([&](auto K) {
            constexpr int k = K;
            cplx<T> acc = in[0];
            sfor<1, R>([&](auto J) {
                constexpr int j = J;
                acc = acc + twiddle<T, long(DIR) * j * k, R>(in[j]);
            });
            t[k] = acc;
        });
    }
}

// ---------------------------------------------------------- register DFT
template <typename T, int L, int DIR>
struct RegDFT {
    // out[k] = sum_{j<L} in[S*j+O] * w_L^{DIR j k}
    template <int S, int O>
    static DGFS_DEV void run(const cplx<T>* in, cplx<T>* out) {
        if constexpr (L == 1) {
            out[0] = in[O];
        } else {
            constexpr int R = smallest_radix(L);
            constexpr int L2 = L / R;
            if constexpr (L2 == 1) {
                cplx<T> t[R];
                sfor<0, R>([&](auto I) { t[I] = in[O + S * I]; });
                butterfly<T, R, DIR>(t);
                sfor<0, R>([&](auto I) { out[I] = t[I]; });
            } else {
                cplx<T> sub[L];
                sfor<0, R>([&](auto Ri) {
                    constexpr int r = Ri;
                    RegDFT<T, L2, DIR>::template run<S * R, O + S * r>(in, sub + r * L2);
                });
                sfor<0, L2>([&](auto Ki) {
                    constexpr int k2 = Ki;
                    cplx<T> t[R];
                    sfor<0, R>([&](auto Ri) {
                        constexpr int r = Ri;
                        t[r] = twiddle<T, long(DIR) * r * k2, L>(sub[r * L2 + k2]);
                    });
                    butterfly<T, R, DIR>(t);
                    sfor<0, R>([&](auto Ri) {
                        constexpr int k1 = Ri;
                        out[k2 + L2 * k1] = t[k1];
                    });
                });
            }
        }
    }
};

// ------------------------------------------------------------- 3-pass engine
// PAD = 1 pads each LDS row by one element (bank spread); PAD = 0 packs the
// cube (smaller, so that more work-groups fit a CU, at the cost of bank
// conflicts in the first pass's stores).
template <typename T, int N_, int S_, int NT_, int DIR_, int PAD = 1>
struct Fft3Engine {
    static constexpr int N = N_;
    static constexpr int S = S_;
    static constexpr int NT = NT_;
    static constexpr int DIR = DIR_;
    static constexpr int NS = N / S;
    static constexpr int NSP = NS + PAD;        // LDS row length
    static constexpr int NPEN = N * NS;         // pass-2 columns == pass-3 pencils
    static constexpr int LDS_ELEMS = N * N * NSP;
    static_assert(N % S == 0, "S must divide N");
    static_assert(NT >= NPEN, "one pass-3 pencil per thread");

    static DGFS_DEV int vidx(int i1, int i2, int j) { return (i1 * N + i2) * NSP + j; }

    // per-work-group twiddles of the residue-r split: wS[s] = w_S^{DIR s r},
    // wN[m] = w_N^{DIR m r}
    struct Split {
        cplx<T> wS[S];
        cplx<T> wN[NS];
        DGFS_DEV void init(int r) {
            sfor<0, S>([&](auto Si) {
                constexpr int s = Si;
                T sn, cs;
                sincospi_t(T(2 * DIR * ((s * r) % S)) / T(S), &sn, &cs);
                wS[s] = mk(cs, sn);
            });
            sfor<0, NS>([&](auto Mi) {
                constexpr int m = Mi;
                T sn, cs;
                sincospi_t(T(2 * DIR * ((m * r) % N)) / T(N), &sn, &cs);
                wN[m] = mk(cs, sn);
            });
        }
        static DGFS_DEV void sincospi_t(float a, float* s, float* c) { sincospif(a, s, c); }
        static DGFS_DEV void sincospi_t(double a, double* s, double* c) { sincospi(a, s, c); }
    };

    // Pass 1: pencils (i1,i2) along a3 -> outputs o3 = S*j + r into V[i1][i2][j].
    // map1(p, i1, i2) assigns pencil p in [0, N*N) ; load(i1, i2, i3) -> cplx<T>
    template <class Map1, class Load>
    static DGFS_DEV void pass1(cplx<T>* V, const Split& sp, Map1&& map1, Load&& load) {
        for (int p = threadIdx.x; p < N * N; p += NT) {
            int i1, i2;
            map1(p, i1, i2);
            cplx<T> a[N];
            sfor<0, N>([&](auto I) { a[I] = load(i1, i2, int(I)); });
            cplx<T> b[NS];
            if constexpr (S == 1) {
                sfor<0, N>([&](auto I) { b[I] = a[I]; });
            } else {
                sfor<0, NS>([&](auto Mi) {
                    constexpr int m = Mi;
                    cplx<T> acc = a[m];
                    sfor<1, S>([&](auto Si) {
                        constexpr int s = Si;
                        acc = acc + a[m + s * NS] * sp.wS[s];
                    });
                    b[m] = (m == 0) ? acc : acc * sp.wN[m];
                });
            }
            cplx<T> o[NS];
            RegDFT<T, NS, DIR>::template run<1, 0>(b, o);
            sfor<0, NS>([&](auto J) { V[vidx(i1, i2, J)] = o[J]; });
        }
    }

    // Pass 2: columns (i1, j) along a2, in place. pre2(i2_const, v) -> v'
    template <class Pre2>
    static DGFS_DEV void pass2(cplx<T>* V, Pre2&& pre2) {
        for (int c = threadIdx.x; c < NPEN; c += NT) {
            const int i1 = c / NS, j = c - (c / NS) * NS;
            cplx<T> a[N], o[N];
            sfor<0, N>([&](auto I) { a[I] = pre2(I, V[vidx(i1, I, j)]); });
            RegDFT<T, N, DIR>::template run<1, 0>(a, o);
            sfor<0, N>([&](auto I) { V[vidx(i1, I, j)] = o[I]; });
        }
    }

    // Pass 3: pencil (o2, j) along a1 for this thread (tid < NPEN).
    // map3(P, o2, j); pre3(i1_const, v); result o[o1] in registers.
    template <class Map3, class Pre3>
    static DGFS_DEV void pass3(const cplx<T>* V, int P, Map3&& map3, Pre3&& pre3,
                               int& o2, int& j, cplx<T>* o) {
        map3(P, o2, j);
        cplx<T> a[N];
        sfor<0, N>([&](auto I) { a[I] = pre3(I, V[vidx(I, o2, j)]); });
        RegDFT<T, N, DIR>::template run<1, 0>(a, o);
    }

    // Whole transform; epi(o2, j, o[]) is called by threads tid < NPEN.
    template <class Map1, class Load, class Pre2, class Map3, class Pre3, class Epi>
    static DGFS_DEV void run(cplx<T>* V, const Split& sp, Map1&& map1, Load&& load,
                             Pre2&& pre2, Map3&& map3, Pre3&& pre3, Epi&& epi) {
        run(V, sp, [] {}, map1, load, pre2, map3, pre3, epi);
    }

    // As above; prologue() runs by every thread after the opening barrier
    // (LDS owned by the caller may be refilled there: readers of the previous
    // call have passed the barrier, readers in this call's epilogue follow
    // the pass-1 barrier).
    template <class Pro, class Map1, class Load, class Pre2, class Map3, class Pre3, class Epi>
    static DGFS_DEV void run(cplx<T>* V, const Split& sp, Pro&& prologue, Map1&& map1,
                             Load&& load, Pre2&& pre2, Map3&& map3, Pre3&& pre3, Epi&& epi) {
        __syncthreads();
        prologue();
        pass1(V, sp, map1, load);
        __syncthreads();
        pass2(V, pre2);
        __syncthreads();
        if (threadIdx.x < NPEN) {
            int o2, j;
            cplx<T> o[N];
            pass3(V, threadIdx.x, map3, pre3, o2, j, o);
            epi(o2, j, o);
        }
    }
};

// ------------------------------------------------------- lane-pair variant
// The partner lane's value (lanes 2m <-> 2m + 1), a DPP quad permutation.
DGFS_DEV float pair_swap(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
}
DGFS_DEV double pair_swap(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(int(unsigned(b)), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(unsigned(b >> 32)), 0xB1, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(hi)) << 32) |
                                          unsigned(lo));
}
template <typename T>
DGFS_DEV cplx<T> pair_swap(cplx<T> v) { return mk(pair_swap(v.x), pair_swap(v.y)); }

// c ? a : b on values (a conditional on two lvalues selects their addresses)
template <typename T>
DGFS_DEV T vsel(bool c, T a, T b) { return c ? a : b; }

// Length-L DFT shared by a lane pair, decimation in time: lane h holds
// x[k] = in[2 k + h], k < L/2, and receives out[k] = X[(L/2) h + k].
template <typename T, int L, int DIR>
DGFS_DEV void pair_dft_dit(const cplx<T>* x, cplx<T>* out, int h) {
    constexpr int H = L / 2;
    cplx<T> y[H];
    RegDFT<T, H, DIR>::template run<1, 0>(x, y);
    sfor<0, H>([&](auto K) {
        constexpr int k = K;
        const cplx<T> w = twiddle<T, long(DIR) * k, L>(y[k]);   // lane 1: w^k Y_1[k]
        // a select of values: `h ? w : y[k]` on the structs would select
        // between their addresses and keep y out of registers
        y[k] = mk(vsel(h, w.x, y[k].x), vsel(h, w.y, y[k].y));
    });
    // X[k] = Y0 + w^k Y1 (lane 0), X[k + H] = Y0 - w^k Y1 (lane 1): one fma
    // with the lane's sign
    const T sg = h ? T(-1) : T(1);
    sfor<0, H>([&](auto K) {
        const cplx<T> r = pair_swap(y[K]);
        out[K] = mk(y[K].x * sg + r.x, y[K].y * sg + r.y);   // contracted to fma
    });
}

// Length-L DFT shared by a lane pair, decimation in frequency: lane h holds
// a[k] = in[(L/2) h + k], k < L/2, and receives out[m] = X[2 m + h].
template <typename T, int L, int DIR>
DGFS_DEV void pair_dft_dif(const cplx<T>* a, cplx<T>* out, int h) {
    constexpr int H = L / 2;
    cplx<T> d[H];
    const T sg = h ? T(-1) : T(1);
    sfor<0, H>([&](auto K) {
        constexpr int k = K;
        const cplx<T> r = pair_swap(a[k]);
        // lane 0: in[k] + in[k + H]; lane 1: in[k] - in[k + H]
        const cplx<T> s = mk(a[k].x * sg + r.x, a[k].y * sg + r.y);
        d[k] = h ? twiddle<T, long(DIR) * k, L>(s) : s;
    });
    RegDFT<T, H, DIR>::template run<1, 0>(d, out);
}

// Fft3Engine's transform with 2 NPEN threads: every pass-2 column and pass-3
// pencil is shared by a lane pair (pair_dft_dit), and pass 1 folds its
// residue split as it loads, so that no thread holds a whole length-N
// vector; twice the waves of Fft3Engine on the same LDS cube.  Pass 3
// leaves lane h of pair P the outputs o1 = (N/2) h + k, k < N/2.
template <typename T, int N_, int S_, int DIR_, int PAD = 1>
struct Fft3Pair : Fft3Engine<T, N_, S_, 2 * N_ * (N_ / S_), DIR_, PAD> {
    using B = Fft3Engine<T, N_, S_, 2 * N_ * (N_ / S_), DIR_, PAD>;
    using B::N;
    using B::S;
    using B::NS;
    using B::NPEN;
    using B::NT;
    using B::DIR;
    using typename B::Split;
    static constexpr int H = N / 2;
    static_assert(N % 2 == 0 && H % 2 == 0, "pair transforms need N divisible by 4");

    // The residue R is a template parameter, so that the split's twiddles
    // w_S^(s R) and w_N^(m R) are compile-time constants (twiddle<> folds the
    // trivial ones).  The fold runs in chunks of CK outputs with a compiler
    // barrier between them, so that at most S CK loads are in flight per thread.
    template <int R, class Map1, class Load>
    static DGFS_DEV void pass1(cplx<T>* V, Map1&& map1, Load&& load) {
        constexpr int CK = std::is_same<T, float>::value && NS % 8 == 0 ? 8 : NS % 4 == 0 ? 4 : 1;
        for (int p = threadIdx.x; p < N * N; p += NT) {
            int i1, i2;
            map1(p, i1, i2);
            cplx<T> b[NS];
            sfor<0, NS / CK>([&](auto Ci) {
                sfor<0, CK>([&](auto Ki) {
                    constexpr int m = int(Ci) * CK + int(Ki);
                    cplx<T> acc = load(i1, i2, m);
                    sfor<1, S>([&](auto Si) {
                        constexpr int s = Si;
                        acc = acc + twiddle<T, long(DIR) * s * R, S>(load(i1, i2, m + s * NS));
                    });
                    b[m] = twiddle<T, long(DIR) * m * R, N>(acc);
                });
                asm volatile("" ::: "memory");
            });
            cplx<T> o[NS];
            RegDFT<T, NS, DIR>::template run<1, 0>(b, o);
            sfor<0, NS>([&](auto J) { V[B::vidx(i1, i2, J)] = o[J]; });
        }
    }

    // The LDS bases below pass through an empty asm statement: the compiler
    // then forms each address from one base per pass instead of keeping
    // H addresses alive across the caller's loop.
    static DGFS_DEV int opaque(int v) {
        asm volatile("" : "+v"(v));
        return v;
    }

    static DGFS_DEV void pass2(cplx<T>* V) {
        const int h = threadIdx.x & 1, c = threadIdx.x >> 1;    // c < NPEN
        const int i1 = c / NS, j = c - (c / NS) * NS;
        const int bi = opaque(B::vidx(i1, h, j));               // + 2 k rows
        const int bo = opaque(B::vidx(i1, H * h, j));           // + k rows
        cplx<T> x[H], o[H];
        sfor<0, H>([&](auto K) { x[K] = V[bi + 2 * int(K) * B::NSP]; });
        pair_dft_dit<T, N, DIR>(x, o, h);
        sfor<0, H>([&](auto K) { V[bo + int(K) * B::NSP] = o[K]; });
    }

    // pencil P = thread / 2 along a1; lane h gets o[k] = output o1 = H h + k
    template <class Map3>
    static DGFS_DEV void pass3(const cplx<T>* V, Map3&& map3, int& o2, int& j, cplx<T>* o) {
        const int h = threadIdx.x & 1;
        map3(int(threadIdx.x >> 1), o2, j);
        const int bi = opaque(B::vidx(h, o2, j));               // + 2 k planes
        cplx<T> x[H];
        sfor<0, H>([&](auto K) { x[K] = V[bi + 2 * int(K) * N * B::NSP]; });
        pair_dft_dit<T, N, DIR>(x, o, h);
    }

    template <int R, class Pro, class Map1, class Load, class Map3, class Epi>
    static DGFS_DEV void run(cplx<T>* V, Pro&& prologue, Map1&& map1, Load&& load, Map3&& map3,
                             Epi&& epi) {
        __syncthreads();
        prologue();
        pass1<R>(V, map1, load);
        __syncthreads();
        pass2(V);
        __syncthreads();
        int o2, j;
        cplx<T> o[H];
        pass3(V, map3, o2, j, o);
        epi(o2, j, o);
    }
};

struct IdentityPre {
    template <class I, class V>
    DGFS_DEV V operator()(I, V v) const { return v; }
};

}  // namespace dgfs
