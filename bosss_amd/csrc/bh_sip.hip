// Matrix-free SIP operator evaluation and on-device assembly for gfx950.
//
// What it computes is the reference's SIP evaluate (SIPLaplace,
// L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250, driven by
// NECQuadratureVolume / NECQuadratureEdge, F/NECQuadratureVolume.cs:377-870,
// F/NECQuadratureEdge.cs:561-1400): volume integral of -nu grad u . grad v,
// interior-face consistency / symmetry / penalty terms and Dirichlet/Neumann
// boundary terms, all by Gauss quadrature of the orthonormal modal basis.
//
// How (MI355X design): one lane per cell.  The reference-element tables
// (basis gradients at volume nodes, traces and normal derivatives at face
// nodes) are identical for every cell, so each table entry is a wave-uniform
// operand fetched with scalar loads; the per-cell work is register-resident
// FMA chains (basis-at-node, flux, flux-to-modes) with no LDS and no atomics.
// Each cell evaluates its own side of all 2D faces (the neighbour's trace is
// recomputed from the neighbour's modes, read through L2), so y is written
// exactly once per apply: HBM traffic = u + y + nu.
#include <array>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <type_traits>

#include "bh_internal.hpp"

namespace bh {

template <typename R>
struct SipArgs {
    int J;
    int n0, n1, n2;
    int nmax;
    const double* __restrict__ width;  // [4][3][nmax]: h, 1/h, sqrt(h), 1/sqrt(h)
    const R* __restrict__ nu;
    const R* __restrict__ G;    // [nqv][D][Np]
    const R* __restrict__ T;    // [2D][nqf][Np]
    const R* __restrict__ Nn;   // [2D][nqf][Np]
    const R* __restrict__ vw;   // [nqv]
    const R* __restrict__ fw;   // [nqf]
    const R* __restrict__ Kv;   // [D][Np][Np] reference stiffness per direction
    double pen;                 // penalty_base * penalty_deg
    int bc_neumann;             // bit per boundary side
    const R* __restrict__ u;
    R* __restrict__ y;
    const double* __restrict__ affine;
    double alpha, beta;
    const R* __restrict__ b0;   // if set: y = b0 + alpha * (A u)  (residual form; beta ignored)
};

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int D>
__device__ __forceinline__ void decode(int c, int n0, int n1, int* idx) {
    if (D == 2) { idx[0] = c / n1; idx[1] = c - idx[0] * n1; idx[2] = 0; }
    else { int s = n0 * n1; idx[2] = c / s; int r = c - idx[2] * s; idx[1] = r / n0; idx[0] = r - idx[1] * n0; }
}
template <int D>
__device__ __forceinline__ int encode(const int* idx, int n0, int n1) {
    return D == 2 ? idx[0] * n1 + idx[1] : idx[0] + idx[1] * n0 + idx[2] * n0 * n1;
}

// cj = (surface area + boundary area) * 0.5 / volume (F/GridData_CellData.cs:666-700)
// (geometry scalar S: double for the FP64 path, float for the FP32 path)
template <int D, typename R, typename S>
__device__ __forceinline__ S cell_cj(const SipArgs<R>& a, const int* idx, const S* h) {
    int nn[3] = {a.n0, a.n1, a.n2};
    S vol = 1, surf = 0, bnd = 0;
    for (int d = 0; d < D; ++d) vol *= h[d];
    for (int d = 0; d < D; ++d) {
        S ar = vol / h[d];
        surf += 2 * ar;
        if (idx[d] == 0) bnd += ar;
        if (idx[d] == nn[d] - 1) bnd += ar;
    }
    return S(0.5) * (surf + bnd) / vol;
}

// Contribution of face f (= 2d + s, outward normal sgn*e_d) of cell P to
// y_P (register array).  uP: modes of P; uQ: modes of the neighbour Q (may be
// all zero), isBnd: boundary face.  Quantities follow ipLaplace.cs:160-250
// written from P's side.
template <int D, int K, typename R, bool SELF, bool NB>
__device__ __forceinline__ void face_term(const SipArgs<R>& a, int f, const R* uP, const R* uQ,
                                          R sdPinv, R sdQinv, R gP, R gQ, R dA, R nuP, R nuQ,
                                          R mu, bool isBnd, R* y) {
    constexpr int Np = binom(K + D, D);
    constexpr int NQF = ipow(K + 1, D - 1);
    const int d = f >> 1, s = f & 1;
    const R sgn = s ? R(1) : R(-1);
    const int fQ = 2 * d + (1 - s);
    const R* __restrict__ TP = a.T + (size_t)f * NQF * Np;
    const R* __restrict__ NP = a.Nn + (size_t)f * NQF * Np;
    const R* __restrict__ TQ = a.T + (size_t)fQ * NQF * Np;
    const R* __restrict__ NQ = a.Nn + (size_t)fQ * NQF * Np;
    const R nuMax = (fabs(nuP) > fabs(nuQ)) ? nuP : nuQ;
#pragma unroll (Np <= 10 ? NQF : 1)
    for (int q = 0; q < NQF; ++q) {
        R vP = 0, dP = 0, vQ = 0, dQ = 0;
        if (SELF) {
#pragma unroll
            for (int n = 0; n < Np; ++n) { vP += TP[q * Np + n] * uP[n]; dP += NP[q * Np + n] * uP[n]; }
        }
        if (NB && !isBnd) {
#pragma unroll
            for (int n = 0; n < Np; ++n) { vQ += TQ[q * Np + n] * uQ[n]; dQ += NQ[q * Np + n] * uQ[n]; }
        }
        vP *= sdPinv;
        dP *= sgn * gP * sdPinv;   // outward normal derivative of u_P
        vQ *= sdQinv;
        dQ *= sgn * gQ * sdQinv;   // derivative of u_Q along P's outward normal
        R cv, cd;                  // coefficients of v_P and of (dv_P/dn)
        if (isBnd) {
            // Dirichlet with g = 0 (the affine part is added separately); Neumann: none
            cv = nuP * dP - R(2) * mu * nuP * vP;
            cd = nuP * vP;
        } else {
            R jump = vP - vQ;
            cv = R(0.5) * (nuP * dP + nuQ * dQ) - mu * nuMax * jump;
            cd = R(0.5) * nuP * jump;
        }
        R wq = a.fw[q] * dA * sdPinv;
        cv *= wq;
        cd *= wq * sgn * gP;
#pragma unroll
        for (int m = 0; m < Np; ++m) y[m] += cv * TP[q * Np + m] + cd * NP[q * Np + m];
    }
}

// Both faces of direction d (f = 2d: '-', f = 2d+1: '+') of cell P, for the apply.
// On the Legendre tensor basis the '+' face rows are the '-' face rows times
// (-1)^{e_d(n)} (values) and -(-1)^{e_d(n)} (normal derivatives, d/dxi_d), so
// P's traces on both faces come from one split into even and odd modes in d and
// the flux of both faces goes back through the '-' rows once.  The flux is the
// same as face_term's, written as linear combinations of the raw traces
// (t = T u, n = N u per side; the scalings sdinv of the two cells, the normal
// sign and 2/h are folded into per-face coefficients, dA sdPinv^2 = 2/h_d).
// uM / uQ: modes of the '-' / '+' neighbour (zero on a boundary side).
// FAST (d is the direction of consecutive cells, one wave = 64 consecutive cells):
// the neighbours' traces are the adjacent lanes' own traces (ds_bpermute), and
// lanes 4q..4q+3 of hv hold, for face node q, the value and derivative traces of
// the cell before the wave on its '+' face and of the cell after it on its '-'
// face; uM / uQ are not read.
template <typename R>
__device__ __forceinline__ R lane_value(R v, int l) {
    if constexpr (sizeof(R) == 8) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    }
}

template <int D, int K, typename R, int d, bool FAST>
__device__ __forceinline__ void face_pair(const SipArgs<R>& a, const R* uP, const R* uM, const R* uQ, const bool* bnd,
                                          const bool* skip, R gP, const R* gQ, const R* rq, R nuP, const R* nuQ,
                                          const R* mu, R* y, R hv, int lane) {
    constexpr int Np = binom(K + D, D);
    constexpr int NQF = ipow(K + 1, D - 1);
    const R* __restrict__ T = a.T + (size_t)(2 * d) * NQF * Np;
    const R* __restrict__ Nr = a.Nn + (size_t)(2 * d) * NQF * Np;
    R c1[2], c2[2], c3[2], c4[2], c5[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const R sgn = s ? R(1) : R(-1);
        const R G = skip[s] ? R(0) : gP;
        const R A = G * nuP * sgn * gP;
        const R nuMax = (fabs(nuP) > fabs(nuQ[s])) ? nuP : nuQ[s];
        c1[s] = (bnd[s] ? R(1) : R(0.5)) * A;
        c3[s] = -G * mu[s] * (bnd[s] ? R(2) * nuP : nuMax);
        c2[s] = bnd[s] ? R(0) : R(0.5) * G * nuQ[s] * sgn * gQ[s] * rq[s];
        c4[s] = bnd[s] ? R(0) : -c3[s] * rq[s];
        c5[s] = bnd[s] ? R(0) : R(-0.5) * A * rq[s];
    }
#pragma unroll (Np <= 4 ? NQF : 1)
    for (int q = 0; q < NQF; ++q) {
        R tE = 0, tO = 0, nE = 0, nO = 0, mE = 0, mO = 0, mnE = 0, mnO = 0, pt = 0, pn = 0;
        static_for<Np>([&](auto NN) {
            constexpr int n = NN;
            constexpr int e = mode_exponent(D, K, n, d);   // d/dxi_d of a mode constant in xi_d is zero
            const R tv = T[q * Np + n];
            if constexpr (e & 1) {
                tO += tv * uP[n];
                if constexpr (!FAST) mO += tv * uM[n];
            } else {
                tE += tv * uP[n];
                if constexpr (!FAST) mE += tv * uM[n];
            }
            if constexpr (!FAST) pt += tv * uQ[n];
            if constexpr (e > 0) {
                const R nv = Nr[q * Np + n];
                if constexpr (e & 1) {
                    nO += nv * uP[n];
                    if constexpr (!FAST) mnO += nv * uM[n];
                } else {
                    nE += nv * uP[n];
                    if constexpr (!FAST) mnE += nv * uM[n];
                }
                if constexpr (!FAST) pn += nv * uQ[n];
            }
        });
        // the '-' neighbour's '+' face (t, n) and the '+' neighbour's '-' face (t, n)
        R mt, mn, qt, qn;
        if constexpr (FAST) {
            const R hTm = lane_value(hv, 4 * q), hNm = lane_value(hv, 4 * q + 1);
            const R hTp = lane_value(hv, 4 * q + 2), hNp = lane_value(hv, 4 * q + 3);
            const int lm = (lane + 63) & 63, lp = (lane + 1) & 63;
            mt = __shfl(lane == 63 ? hTm : tE - tO, lm, 64);   // lane 0 receives the cell before the wave
            mn = __shfl(lane == 63 ? hNm : nO - nE, lm, 64);
            qt = __shfl(lane == 0 ? hTp : tE + tO, lp, 64);    // lane 63 receives the cell after it
            qn = __shfl(lane == 0 ? hNp : nE + nO, lp, 64);
        } else {
            mt = mE - mO;
            mn = mnO - mnE;
            qt = pt;
            qn = pn;
        }
        // '-' face: P's t, n = tE + tO, nE + nO
        const R cvm = c1[0] * (nE + nO) + c2[0] * mn + c3[0] * (tE + tO) + c4[0] * mt;
        const R cdm = c1[0] * (tE + tO) + c5[0] * mt;
        // '+' face: P's t, n = tE - tO, nO - nE
        const R cvp = c1[1] * (nO - nE) + c2[1] * qn + c3[1] * (tE - tO) + c4[1] * qt;
        const R cdp = c1[1] * (tE - tO) + c5[1] * qt;
        const R w = a.fw[q];
        const R TEc = w * (cvm + cvp), TOc = w * (cvm - cvp), NEc = w * (cdm - cdp), NOc = w * (cdm + cdp);
        static_for<Np>([&](auto MM) {
            constexpr int m = MM;
            constexpr int e = mode_exponent(D, K, m, d);
            y[m] = fma(T[q * Np + m], (e & 1) ? TOc : TEc, y[m]);
            if constexpr (e > 0) y[m] = fma(Nr[q * Np + m], (e & 1) ? NOc : NEc, y[m]);
        });
    }
}

// Modal form of face_pair.  On an axis-aligned box the face rule of k + 1 Gauss
// points per direction integrates every SIP face product exactly, and on the
// face x_d = -1 / +1 the basis function of exponents (e', j) (e' the exponents
// across the face, j the one along d) is p_j(-1 / +1) Phi_e'; its d-derivative
// is p_j'(-1 / +1) Phi_e', with p_j = sqrt(j + 1/2) P_j, so p_j(+-1) = (+-1)^j s_j
// and p_j'(+-1) = (+-1)^(j+1) s_j j(j+1)/2.  The Phi_e' are orthonormal on the
// face, so the node sums of face_pair reduce to sums over e': per face mode the
// traces are A = sum_j p_j(+-1) u_(e',j) and B = sum_j p_j'(+-1) u_(e',j), the
// flux terms cvm, cdm, cvp, cdp become per-mode scalars, and
// y_(e',j) += p_j(-1) CVM + p_j(+1) CVP + p_j'(-1) CDM + p_j'(+1) CDP.
// The result equals the quadrature up to rounding, with about 4 Np multiply-adds
// per side instead of 4 Np per face node.  A face mode is addressed by its root:
// the mode with the same e' and j = 0.

// The modal face terms of one side of direction d: s = 0 is the '-' face (neighbour
// M, its '+' face), s = 1 the '+' face (neighbour Q, its '-' face).  The coefficients
// c1..c5 are those of face_pair.  The two sides are independent, so the kernel runs
// them one after the other with one neighbour's coefficients live at a time.
template <int D, int K, typename R, int d, int s>
__device__ __forceinline__ void face_side_modal(R* uP, R* un, bool bnd, bool skip, R gP, R gQ, R rq,
                                                R nuP, R nuQ, R mu, R* y) {
    constexpr int Np = binom(K + D, D);
    static_assert(K <= 6, "sqrt_half covers j <= 6");
    const R sgn = s ? R(1) : R(-1);
    const R G = skip ? R(0) : gP;
    const R A = G * nuP * sgn * gP;
    const R nuMax = (fabs(nuP) > fabs(nuQ)) ? nuP : nuQ;
    const R c1 = (bnd ? R(1) : R(0.5)) * A;
    const R c3 = -G * mu * (bnd ? R(2) * nuP : nuMax);
    const R c2 = bnd ? R(0) : R(0.5) * G * nuQ * sgn * gQ * rq;
    const R c4 = bnd ? R(0) : -c3 * rq;
    const R c5 = bnd ? R(0) : R(-0.5) * A * rq;
    static_for<Np>([&](auto FF) {
        constexpr int f = FF;
        if constexpr (mode_exponent(D, K, f, d) == 0) {
            // P's value / d-derivative trace on face s, the neighbour's on the same face
            R Ao = 0, Bo = 0, An = 0, Bn = 0;
            static_for<Np>([&](auto NN) {
                constexpr int n = NN;
                if constexpr (face_root(D, K, n, d) == f) {
                    constexpr int j = mode_exponent(D, K, n, d);
                    constexpr double sv = sqrt_half(j), sd = sqrt_half(j) * (j * (j + 1) / 2);
                    constexpr double odd = (j & 1) ? -1.0 : 1.0;
                    // p_j(+1) = s_j, p_j'(+1) = sd_j; p_j(-1) = (-1)^j s_j, p_j'(-1) = -(-1)^j sd_j
                    constexpr double vo = s ? sv : odd * sv, dov = s ? sd : -odd * sd;   // own side
                    constexpr double vn = s ? odd * sv : sv, dnv = s ? -odd * sd : sd;   // neighbour's side
                    asm volatile("" : "+v"(uP[n]), "+v"(un[n]));   // in place: no copies
                    Ao = fma(R(vo), uP[n], Ao);
                    An = fma(R(vn), un[n], An);
                    if constexpr (j > 0) {
                        Bo = fma(R(dov), uP[n], Bo);
                        Bn = fma(R(dnv), un[n], Bn);
                    }
                }
            });
            const R CV = c1 * Bo + c2 * Bn + c3 * Ao + c4 * An;
            const R CD = c1 * Ao + c5 * An;
            static_for<Np>([&](auto MM) {
                constexpr int m = MM;
                if constexpr (face_root(D, K, m, d) == f) {
                    constexpr int j = mode_exponent(D, K, m, d);
                    constexpr double sv = sqrt_half(j), sd = sqrt_half(j) * (j * (j + 1) / 2);
                    constexpr double odd = (j & 1) ? -1.0 : 1.0;
                    constexpr double vo = s ? sv : odd * sv, dov = s ? sd : -odd * sd;
                    y[m] = fma(R(vo), CV, y[m]);
                    if constexpr (j > 0) y[m] = fma(R(dov), CD, y[m]);
                    asm volatile("" : "+v"(y[m]));
                }
            });
        }
    });
}

// Volume integral -nu grad u . grad v of an axis-aligned box cell: the metric
// is diagonal and constant, so the Gauss sum over the volume nodes factorises
// into the per-direction reference stiffness Kv[d] = sum_q w_q G_d(q) G_d(q)^T
// (computed once from the same rule): y += sum_d (-nu 4/h_d^2) Kv[d] u.
// On the orthonormal Legendre tensor basis Kv[d][m][n] vanishes unless the
// exponents agree in every other direction and both are >= 1 and of equal
// parity in d (the rule is exact for these products): only those entries are
// summed.
__host__ __device__ constexpr bool kv_nonzero(int D, int k, int m, int n, int d) {
    for (int dd = 0; dd < D; ++dd) {
        const int em = mode_exponent(D, k, m, dd), en = mode_exponent(D, k, n, dd);
        if (dd == d) {
            if (em == 0 || en == 0 || ((em + en) & 1)) return false;
        } else if (em != en) {
            return false;
        }
    }
    return true;
}

template <int D, int K, typename R>
__device__ __forceinline__ void volume_term(const SipArgs<R>& a, const R* uP, const R* s2, R nuP, R* y) {
    constexpr int Np = binom(K + D, D);
    static_for<D>([&](auto DD) {
        constexpr int d = DD;
        const R* __restrict__ Kd = a.Kv + (size_t)d * Np * Np;
        const R cd = -nuP * s2[d];
        static_for<Np>([&](auto MM) {
            constexpr int m = MM;
            R acc = 0;
            bool any = false;
            static_for<Np>([&](auto NN) {
                constexpr int n = NN;
                if constexpr (kv_nonzero(D, K, m, n, d)) {
                    acc += Kd[m * Np + n] * uP[n];
                    any = true;
                }
            });
            if (any) y[m] += cd * acc;
        });
    });
}

// The same volume term with Kv from its closed form: on the orthonormal Legendre
// basis p_j = sqrt(j + 1/2) P_j, int_{-1}^{1} p_a' p_b' = sqrt((a + 1/2)(b + 1/2)) m (m + 1)
// with m = min(a, b) for a, b >= 1 and a + b even (from P_n' = sum over k = n-1, n-3, ...
// of (2k + 1) P_k), and the k + 1 point Gauss rule integrates these products exactly,
// so Kv[d][m][n] is that value when the exponents agree in the other directions.
// The entries become compile-time constants: no table loads, no registers held.
__host__ __device__ constexpr double kv_closed(int a, int b) {
    const int m = a < b ? a : b;
    return sqrt_half(a) * sqrt_half(b) * (double)(m * (m + 1));
}
template <int D, int K, typename R>
__device__ __forceinline__ void volume_term_cf(R* uP, const R* s2, R nuP, R* y) {
    constexpr int Np = binom(K + D, D);
    static_for<D>([&](auto DD) {
        constexpr int d = DD;
        const R cd = -nuP * s2[d];
        static_for<Np>([&](auto MM) {
            constexpr int m = MM;
            R acc = 0;
            bool any = false;
            static_for<Np>([&](auto NN) {
                constexpr int n = NN;
                if constexpr (kv_nonzero(D, K, m, n, d)) {
                    constexpr double kv = kv_closed(mode_exponent(D, K, m, d), mode_exponent(D, K, n, d));
                    acc = fma(R(kv), uP[n], acc);
                    any = true;
                }
            });
            if (any) {
                y[m] = fma(cd, acc, y[m]);
                asm volatile("" : "+v"(y[m]));
            }
        });
    });
}

// geometry of cell with logical index idx
template <int D, typename R, typename S>
__device__ __forceinline__ void cell_geom(const SipArgs<R>& a, const int* idx, S* h) {
    for (int d = 0; d < D; ++d) h[d] = S(a.width[d * a.nmax + idx[d]]);
}

// Writes the results of a wave's 64 consecutive cells (lane l: cell c0 + l, its Np
// values y; the first nvalid lanes valid) as 16-byte chunks that consecutive lanes
// write to consecutive addresses: each lane stores its values in a per-wave LDS buffer
// (cell stride Np + 1, so the lanes' rows fall into different banks), and the wave
// reads them back in chunk order.  One lane per cell writing its own Np values would
// issue Np stores per wave that each touch 64 scattered rows.
// out = alpha * (y + affine) + (b0 ? b0 : beta * out), element by element.
typedef double nt_double2 __attribute__((ext_vector_type(2)));
typedef float nt_float4 __attribute__((ext_vector_type(4)));

template <int Np, typename R>
__device__ __forceinline__ void store_cells_coalesced(R* __restrict__ buf, const R* y, int lane, size_t c0, int nvalid,
                                                      const SipArgs<R>& a) {
#pragma unroll
    for (int n = 0; n < Np; ++n) buf[lane * (Np + 1) + n] = y[n];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int CH = 16 / sizeof(R);             // elements per chunk
    constexpr int NCH = (64 * Np + CH - 1) / CH;   // chunks per wave
    const int nel = nvalid * Np;
    R* __restrict__ yp = a.y + c0 * Np;
    const R* __restrict__ bp = a.b0 ? a.b0 + c0 * Np : nullptr;
    const double* __restrict__ af = a.affine ? a.affine + c0 * Np : nullptr;
    const R al = R(a.alpha), be = R(a.beta);
    const bool full_chunks = ((c0 * Np) % CH) == 0;   // chunk-aligned block start
#pragma unroll
    for (int i = 0; i < (NCH + 63) / 64; ++i) {
        const int q = lane + 64 * i;
        const int e0 = q * CH;
        if (e0 >= nel) break;
        R v[CH];
#pragma unroll
        for (int t = 0; t < CH; ++t) {
            const int e = e0 + t;
            const int cell = e / Np, m = e - cell * Np;
            v[t] = e < nel ? buf[cell * (Np + 1) + m] : R(0);
        }
        if (af) {
#pragma unroll
            for (int t = 0; t < CH; ++t) if (e0 + t < nel) v[t] += R(af[e0 + t]);
        }
        if (full_chunks && e0 + CH <= nel) {
            using V = std::conditional_t<sizeof(R) == 8, double2, float4>;
            V o;
            R* ov = reinterpret_cast<R*>(&o);
            if (bp) {
                const V b = *reinterpret_cast<const V*>(bp + e0);
                const R* bv = reinterpret_cast<const R*>(&b);
#pragma unroll
                for (int t = 0; t < CH; ++t) ov[t] = bv[t] + al * v[t];
            } else if (a.beta == 0.0) {
#pragma unroll
                for (int t = 0; t < CH; ++t) ov[t] = al * v[t];
            } else {
                const V b = *reinterpret_cast<const V*>(yp + e0);
                const R* bv = reinterpret_cast<const R*>(&b);
#pragma unroll
                for (int t = 0; t < CH; ++t) ov[t] = be * bv[t] + al * v[t];
            }
            // streaming store: the results are not read again by this kernel, so they
            // should not displace the neighbour rows other waves still read from L2
            using NV = std::conditional_t<sizeof(R) == 8, nt_double2, nt_float4>;
            NV nv;
            __builtin_memcpy(&nv, &o, sizeof(NV));
            __builtin_nontemporal_store(nv, reinterpret_cast<NV*>(yp + e0));
        } else {
#pragma unroll
            for (int t = 0; t < CH; ++t) {
                const int e = e0 + t;
                if (e >= nel) break;
                yp[e] = bp ? bp[e] + al * v[t] : (a.beta == 0.0 ? al * v[t] : be * yp[e] + al * v[t]);
            }
        }
    }
}

// MODAL: every direction loads its neighbours' coefficients and uses the modal
// face form (face_side_modal); otherwise the face rule is evaluated node by node.
template <int D, int K, typename R, bool MODAL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void sip_apply_kernel(SipArgs<R> a) {
    constexpr int Np = binom(K + D, D);
    constexpr int NQF = ipow(K + 1, D - 1);
    constexpr int DF = D == 2 ? 1 : 0;       // direction of consecutive cells
    const int c0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63);   // first cell of the wave
    const int lane = threadIdx.x & 63;
    const int craw = c0 + lane;
    const bool valid = craw < a.J;
    const int c = valid ? craw : a.J - 1;   // every lane stays active for the lane exchanges
    int idx[3];
    decode<D>(c, a.n0, a.n1, idx);
    const int nn[3] = {a.n0, a.n1, a.n2};
    R uP[Np], y[Np];
    const R* up = a.u + (size_t)c * Np;
#pragma unroll
    for (int n = 0; n < Np; ++n) { uP[n] = up[n]; y[n] = 0; }
    const R nuP = a.nu[c];
    // reciprocal widths once; cj = sum_d (1 + boundary faces_d / 2) / h_d (the cell_cj formula),
    // and the neighbour's cj differs from P's only in direction d
    R ih[3] = {1, 1, 1}, s2[D], cjP = 0;
    for (int d = 0; d < D; ++d) {
        ih[d] = R(a.width[(3 + d) * a.nmax + idx[d]]);
        s2[d] = R(4) * ih[d] * ih[d];
        cjP += (R(1) + R(0.5) * R((idx[d] == 0) + (idx[d] == nn[d] - 1))) * ih[d];
    }
    if constexpr (MODAL) volume_term_cf<D, K, R>(uP, s2, nuP, y);
    else volume_term<D, K, R>(a, uP, s2, nuP, y);
    const R pen = R(a.pen);
    static_for<D>([&](auto DD) {
        constexpr int d = DD;
        const R ihd = ih[d];
        if constexpr (MODAL) {
            // modal faces, one side at a time.  A boundary side reads the cell itself as its
            // "neighbour": the flux coefficients of the neighbour's traces (c2, c4, c5) are zero
            // there, so any finite values do, and the loads need no branch or zero fill.
            static_for<2>([&](auto SS) {
                constexpr int s = SS;
                const bool bnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
                const bool skip = bnd && ((a.bc_neumann >> (2 * d + s)) & 1);
                int nidx[3] = {idx[0], idx[1], idx[2]};
                nidx[d] += bnd ? 0 : (s ? 1 : -1);
                const int cq = encode<D>(nidx, a.n0, a.n1);
                const R ihq = R(a.width[(3 + d) * a.nmax + nidx[d]]);
                const R bP = R(1) + R(0.5) * R((idx[d] == 0) + (idx[d] == nn[d] - 1));
                const R bQ = R(1) + R(0.5) * R((nidx[d] == 0) + (nidx[d] == nn[d] - 1));
                const R cjQ = cjP - bP * ihd + bQ * ihq;   // = cjP on a boundary side
                const R gQ = R(2) * ihq;
                const R rq = R(a.width[(6 + d) * a.nmax + idx[d]] * a.width[(9 + d) * a.nmax + nidx[d]]);
                const R nuQ = a.nu[cq];
                const R mu = pen * (cjP > cjQ ? cjP : cjQ);
                R un[Np];
                const R* uq = a.u + (size_t)cq * Np;
#pragma unroll
                for (int n = 0; n < Np; ++n) un[n] = uq[n];
                face_side_modal<D, K, R, d, s>(uP, un, bnd, skip, R(2) * ihd, gQ, rq, nuP, nuQ, mu, y);
            });
            return;
        }
        constexpr bool FAST = d == DF && NQF <= 16 && Np > 4;   // at k = 1 the exchange costs what it saves
        bool bnd[2], skip[2];
        R gQ[2], rq[2], nuQ[2], mu[2];
        R uM[FAST ? 1 : Np], uQ[FAST ? 1 : Np];
        R hv = 0;
        if constexpr (FAST) {
            // traces of the cells before and after the wave, node q = lane / 4:
            // kind 0 / 1: value / derivative of cell c0 - 1 on its '+' face (the '-' face
            // rows times (-1)^e, and times -(-1)^e for the derivative); kind 2 / 3: value /
            // derivative of cell c0 + 64 on its '-' face
            const int hq = lane >> 2, kind = lane & 3;
            if (hq < NQF) {
                const int hc = kind < 2 ? max(c0 - 1, 0) : min(c0 + 64, a.J - 1);
                const R* __restrict__ row = (kind & 1 ? a.Nn : a.T) + ((size_t)(2 * d) * NQF + hq) * Np;
                const R* __restrict__ uh = a.u + (size_t)hc * Np;
                static_for<Np>([&](auto NN) {
                    constexpr int n = NN;
                    constexpr int e = mode_exponent(D, K, n, d);
                    const R v = row[n] * uh[n];
                    hv += (kind < 2 && (e & 1)) ? -v : v;
                });
                if (kind == 1) hv = -hv;
            }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            R* un = s ? uQ : uM;
            bnd[s] = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
            skip[s] = bnd[s] && ((a.bc_neumann >> (2 * d + s)) & 1);
            if (bnd[s]) {
                gQ[s] = 0; rq[s] = 0; nuQ[s] = nuP; mu[s] = pen * cjP;
                if constexpr (!FAST) {
#pragma unroll
                    for (int n = 0; n < Np; ++n) un[n] = 0;
                }
            } else {
                int nidx[3] = {idx[0], idx[1], idx[2]};
                nidx[d] += s ? 1 : -1;
                const int cq = encode<D>(nidx, a.n0, a.n1);
                const R ihq = R(a.width[(3 + d) * a.nmax + nidx[d]]);
                const R bP = R(1) + R(0.5) * R((idx[d] == 0) + (idx[d] == nn[d] - 1));
                const R bQ = R(1) + R(0.5) * R((nidx[d] == 0) + (nidx[d] == nn[d] - 1));
                const R cjQ = cjP - bP * ihd + bQ * ihq;
                gQ[s] = R(2) * ihq;
                rq[s] = R(a.width[(6 + d) * a.nmax + idx[d]] * a.width[(9 + d) * a.nmax + nidx[d]]);   // sdQinv / sdPinv
                nuQ[s] = a.nu[cq];
                mu[s] = pen * (cjP > cjQ ? cjP : cjQ);
                if constexpr (!FAST) {
                    const R* uq = a.u + (size_t)cq * Np;
#pragma unroll
                    for (int n = 0; n < Np; ++n) un[n] = uq[n];
                }
            }
        }
        face_pair<D, K, R, d, FAST>(a, uP, uM, uQ, bnd, skip, R(2) * ihd, gQ, rq, nuP, nuQ, mu, y, hv, lane);
    });
    __shared__ R sbuf[4][64 * (Np + 1)];
    const int nvalid = a.J - c0 < 64 ? a.J - c0 : 64;
    store_cells_coalesced<Np, R>(sbuf[threadIdx.x >> 6], y, lane, (size_t)c0, nvalid, a);
}

// ------------------------------------------------------------------ tiled modal apply
// The modal face form with the neighbours' traces exchanged instead of reloaded.
// Direction roles: DF, the direction of consecutive cells, runs along the 64 lanes
// of a wave; DT runs along the W waves of a workgroup (one row of up to 64 cells
// each); DZ (3D) is the remaining direction.  Per face mode e' a face needs only
// the value and d-derivative traces (A, B) of both cells, so:
//  * DF: every lane forms its own traces on both faces and takes the neighbours'
//    from the adjacent lanes (DPP wave shifts); the cells just before and after
//    the wave (rows longer than 64 cells) are traced once, across its lanes;
//  * DT: every wave writes its traces on both DT faces to LDS; after one workgroup
//    barrier, a wave reads the adjacent waves' traces, and only the first and last
//    wave of the workgroup load a neighbour row from memory;
//  * DZ: the neighbours' modes are loaded, as in sip_apply_kernel.
// Workgroups are mapped to tiles XCD-contiguously (tile = (b mod 8) * per_xcd +
// b / 8, the hardware dispatching workgroup b to XCD b mod 8), so the DZ
// neighbour rows of a tile are read by the same XCD's L2 shortly before or after.
// Same integrals and the same flux as face_side_modal; results equal
// sip_apply_kernel<.., MODAL = true> up to rounding.

// index of face mode f (a root: exponent 0 in d) among the roots of direction d
__host__ __device__ constexpr int face_slot(int D, int k, int f, int d) {
    int r = 0;
    for (int m = 0; m < f; ++m) if (mode_exponent(D, k, m, d) == 0) ++r;
    return r;
}

// value / d-derivative traces of face mode f on face `side` (0: '-', 1: '+') of
// direction d: A = sum_j p_j(+-1) u_(e',j), B = sum_j p_j'(+-1) u_(e',j)
template <int D, int K, typename R, int d, int side, int f>
__device__ __forceinline__ void modal_trace(R* u, R& A, R& B) {
    constexpr int Np = binom(K + D, D);
    A = 0;
    B = 0;
    static_for<Np>([&](auto NN) {
        constexpr int n = NN;
        if constexpr (face_root(D, K, n, d) == f) {
            constexpr int j = mode_exponent(D, K, n, d);
            constexpr double sv = sqrt_half(j), sd = sqrt_half(j) * (j * (j + 1) / 2);
            constexpr double odd = (j & 1) ? -1.0 : 1.0;
            constexpr double cv = side ? sv : odd * sv, cdv = side ? sd : -odd * sd;
            asm volatile("" : "+v"(u[n]));   // one face mode's inputs at a time, pinned in place
            A = fma(R(cv), u[n], A);
            if constexpr (j > 0) B = fma(R(cdv), u[n], B);
        }
    });
}

// y_(e',j) += p_j(+-1) CV + p_j'(+-1) CD on face `side` of direction d
template <int D, int K, typename R, int d, int side, int f>
__device__ __forceinline__ void modal_scatter(R CV, R CD, R* y) {
    constexpr int Np = binom(K + D, D);
    static_for<Np>([&](auto MM) {
        constexpr int m = MM;
        if constexpr (face_root(D, K, m, d) == f) {
            constexpr int j = mode_exponent(D, K, m, d);
            constexpr double sv = sqrt_half(j), sd = sqrt_half(j) * (j * (j + 1) / 2);
            constexpr double odd = (j & 1) ? -1.0 : 1.0;
            constexpr double cv = side ? sv : odd * sv, cdv = side ? sd : -odd * sd;
            y[m] = fma(R(cv), CV, y[m]);
            if constexpr (j > 0) y[m] = fma(R(cdv), CD, y[m]);
            asm volatile("" : "+v"(y[m]));
        }
    });
}

// SIP flux coefficients of one side (those of face_pair / face_side_modal):
// CV = c1 Bo + c2 Bn + c3 Ao + c4 An, CD = c1 Ao + c5 An (o: own, n: neighbour)
template <typename R>
struct SideCoef {
    R c1, c2, c3, c4, c5;
};
template <typename R, int s>
__device__ __forceinline__ SideCoef<R> side_coef(bool bnd, bool skip, R gP, R gQ, R rq, R nuP, R nuQ, R mu) {
    const R sgn = s ? R(1) : R(-1);
    const R G = skip ? R(0) : gP;
    const R A = G * nuP * sgn * gP;
    const R nuMax = (fabs(nuP) > fabs(nuQ)) ? nuP : nuQ;
    SideCoef<R> c;
    c.c1 = (bnd ? R(1) : R(0.5)) * A;
    c.c3 = -G * mu * (bnd ? R(2) * nuP : nuMax);
    c.c2 = bnd ? R(0) : R(0.5) * G * nuQ * sgn * gQ * rq;
    c.c4 = bnd ? R(0) : -c.c3 * rq;
    c.c5 = bnd ? R(0) : R(-0.5) * A * rq;
    return c;
}

// Face data of side s of direction d of the cell idx: boundary / Neumann flags and,
// for an interior face, the neighbour's index, 2/h, sqrt-width ratio and penalty.
template <int D, typename R, int d, int s>
struct SideGeom {
    bool bnd, skip;
    int cq;
    R gQ, rq, mu;
};
// A boundary side gets the cell itself as its neighbour (cq = own index, the penalty of
// the own cell): the flux coefficients that multiply the neighbour's traces are zero
// there, so reading finite values of the cell itself needs no branch or zero fill.
template <int D, typename R, int d, int s>
__device__ __forceinline__ SideGeom<D, R, d, s> side_geom(const SipArgs<R>& a, const int* idx, const int* nn, R cjP,
                                                          R ihd, R pen) {
    SideGeom<D, R, d, s> g;
    g.bnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
    g.skip = g.bnd && ((a.bc_neumann >> (2 * d + s)) & 1);
    int nidx[3] = {idx[0], idx[1], idx[2]};
    nidx[d] += g.bnd ? 0 : (s ? 1 : -1);
    g.cq = encode<D>(nidx, a.n0, a.n1);
    const R ihq = R(a.width[(3 + d) * a.nmax + nidx[d]]);
    const R bP = R(1) + R(0.5) * R((idx[d] == 0) + (idx[d] == nn[d] - 1));
    const R bQ = R(1) + R(0.5) * R((nidx[d] == 0) + (nidx[d] == nn[d] - 1));
    const R cjQ = cjP - bP * ihd + bQ * ihq;   // = cjP on a boundary side
    g.gQ = R(2) * ihq;
    g.rq = R(a.width[(6 + d) * a.nmax + idx[d]] * a.width[(9 + d) * a.nmax + nidx[d]]);
    g.mu = pen * (cjP > cjQ ? cjP : cjQ);
    return g;
}

// whole-wave lane shift of a value (DPP): WF_SR1 (0x138) gives lane l the value of
// lane l - 1, WF_SL1 (0x130) that of lane l + 1; the lane without a source keeps old
template <int CTRL, typename R>
__device__ __forceinline__ R wave_shift(R old, R src) {
    if constexpr (sizeof(R) == 8) {
        const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
        const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xf, 0xf, false);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    } else {
        return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xf, 0xf, false));
    }
}

// One side of direction d from the neighbour's traces: nb(f, fs, An, Bn) supplies them.
template <int D, int K, typename R, int d, int s, typename NB>
__device__ __forceinline__ void face_side_traces(R* uP, const SideCoef<R>& c, NB&& nb, R* y) {
    constexpr int Np = binom(K + D, D);
    static_for<Np>([&](auto FF) {
        constexpr int f = FF;
        if constexpr (mode_exponent(D, K, f, d) == 0) {
            constexpr int fs = face_slot(D, K, f, d);
            R Ao, Bo, An, Bn;
            modal_trace<D, K, R, d, s, f>(uP, Ao, Bo);
            nb(std::integral_constant<int, f>{}, std::integral_constant<int, fs>{}, An, Bn);
            asm volatile("" : "+v"(Ao), "+v"(Bo), "+v"(An), "+v"(Bn));
            const R CV = c.c1 * Bo + c.c2 * Bn + c.c3 * Ao + c.c4 * An;
            const R CD = c.c1 * Ao + c.c5 * An;
            modal_scatter<D, K, R, d, s, f>(CV, CD, y);
        }
    });
}

template <int D, int K, typename R, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(3))) void sip_apply_tile_kernel(SipArgs<R> a, int nseg, int nrb, int ntiles, int per_xcd) {
    constexpr int Np = binom(K + D, D);
    constexpr int NFM = binom(K + D - 1, D - 1);   // face modes
    constexpr int DF = D == 2 ? 1 : 0;             // consecutive cells: lanes
    constexpr int DT = D == 2 ? 0 : 1;             // rows of the workgroup: LDS
    // one face's traces per wave at a time (A, B per face mode), then the staged results:
    // at 64^3 k = 3 FP64 43 KB per workgroup of 4 waves, 3 workgroups per CU
    constexpr int TRW = 2 * NFM > Np + 1 ? 2 * NFM : Np + 1;
    __shared__ R tr[W][TRW][64];
    const int t = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (t >= ntiles) return;                       // whole workgroup, before any barrier
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int seg = t % nseg, rest = t / nseg;
    const int rb = rest % nrb, z = rest / nrb;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int x0 = seg * 64, x = x0 + lane, r = rb * W + w;
    const bool valid = x < nn[DF] && r < nn[DT];
    int idx[3] = {0, 0, 0};
    idx[DF] = x < nn[DF] ? x : nn[DF] - 1;         // every lane stays active for the exchanges
    idx[DT] = r < nn[DT] ? r : nn[DT] - 1;
    if (D == 3) idx[2] = z;
    const int c = encode<D>(idx, a.n0, a.n1);
    R uP[Np], y[Np];
    const R* up = a.u + (size_t)c * Np;
#pragma unroll
    for (int n = 0; n < Np; ++n) { uP[n] = up[n]; y[n] = 0; }
    const R nuP = a.nu[c];
    // Small cells (k = 1, and k = 2 in FP32): the neighbour rows that come from memory
    // (both DZ sides; the DT side of the workgroup's first or last wave) are loaded
    // here, so their latency overlaps the DF and DT work instead of following it.  A
    // boundary side loads the cell itself (finite, multiplied by zero flux weights).
    constexpr bool PRE = Np * (int)sizeof(R) <= 40;
    R zq[2][PRE && D == 3 ? Np : 1], tq[PRE ? Np : 1];
    if constexpr (PRE) {
        if constexpr (D == 3) {
            static_for<2>([&](auto SS) {
                constexpr int s = SS;
                int nidx[3] = {idx[0], idx[1], idx[2]};
                nidx[2] += (s ? idx[2] == nn[2] - 1 : idx[2] == 0) ? 0 : (s ? 1 : -1);
                const R* uq = a.u + (size_t)encode<D>(nidx, a.n0, a.n1) * Np;
#pragma unroll
                for (int n = 0; n < Np; ++n) zq[s][n] = uq[n];
            });
        }
        if (w == 0 || w == W - 1) {   // wave-uniform
            const int s = w == 0 ? 0 : 1;
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[DT] += (s ? idx[DT] == nn[DT] - 1 : idx[DT] == 0) ? 0 : (s ? 1 : -1);
            const R* uq = a.u + (size_t)encode<D>(nidx, a.n0, a.n1) * Np;
#pragma unroll
            for (int n = 0; n < Np; ++n) tq[n] = uq[n];
        }
    }
    R ih[3] = {1, 1, 1}, s2[D], cjP = 0;
    for (int d = 0; d < D; ++d) {
        ih[d] = R(a.width[(3 + d) * a.nmax + idx[d]]);
        s2[d] = R(4) * ih[d] * ih[d];
        cjP += (R(1) + R(0.5) * R((idx[d] == 0) + (idx[d] == nn[d] - 1))) * ih[d];
    }
    const R pen = R(a.pen);
    // 1. own traces on the '+' DT face -> LDS (the '-' face's follow after the first DT side)
    auto put_traces = [&](auto SS) {
        constexpr int s = SS;
        static_for<Np>([&](auto FF) {
            constexpr int f = FF;
            if constexpr (mode_exponent(D, K, f, DT) == 0) {
                constexpr int fs = face_slot(D, K, f, DT);
                R A, B;
                modal_trace<D, K, R, DT, s, f>(uP, A, B);
                tr[w][2 * fs][lane] = A;
                tr[w][2 * fs + 1][lane] = B;
            }
        });
    };
    put_traces(std::integral_constant<int, 1>{});
    // 2. volume
    volume_term_cf<D, K, R>(uP, s2, nuP, y);
    // 3. DF faces from the adjacent lanes
    {
        constexpr int d = DF;
        const R ihd = ih[d];
        const auto gm = side_geom<D, R, d, 0>(a, idx, nn, cjP, ihd, pen);
        const auto gp = side_geom<D, R, d, 1>(a, idx, nn, cjP, ihd, pen);
        const R nuM = a.nu[gm.cq], nuQ = a.nu[gp.cq];
        const SideCoef<R> cm = side_coef<R, 0>(gm.bnd, gm.skip, R(2) * ihd, gm.gQ, gm.rq, nuP, nuM, gm.mu);
        const SideCoef<R> cp = side_coef<R, 1>(gp.bnd, gp.skip, R(2) * ihd, gp.gQ, gp.rq, nuP, nuQ, gp.mu);
        // traces of the cells before and after the wave: lanes 2 fs + {0, 1} (A, B of face
        // mode slot fs) for the cell before (its '+' face), lanes 32 + 2 fs + {0, 1} for the
        // cell after (its '-' face)
        const bool halo = x0 > 0 || x0 + 64 < nn[d];   // wave-uniform
        R hv = 0;
        if (halo) {
            const int hl = lane & 31;
            const bool after = lane >= 32;
            const int hx = after ? x0 + 64 : x0 - 1;
            if (hl < 2 * NFM && hx >= 0 && hx < nn[d]) {
                int hidx[3] = {idx[0], idx[1], idx[2]};
                hidx[d] = hx;
                const R* uh = a.u + (size_t)encode<D>(hidx, a.n0, a.n1) * Np;
                const int fsl = hl >> 1, kind = hl & 1;
                static_for<Np>([&](auto NN) {
                    constexpr int n = NN;
                    constexpr int j = mode_exponent(D, K, n, d);
                    constexpr int fsn = face_slot(D, K, face_root(D, K, n, d), d);
                    constexpr double sv = sqrt_half(j), sd = sqrt_half(j) * (j * (j + 1) / 2);
                    constexpr double odd = (j & 1) ? -1.0 : 1.0;
                    if (fsn == fsl && (kind == 0 || j > 0)) {
                        const R cf = kind ? (after ? R(-odd * sd) : R(sd)) : (after ? R(odd * sv) : R(sv));
                        hv = fma(cf, uh[n], hv);
                    }
                });
            }
        }
        static_for<Np>([&](auto FF) {
            constexpr int f = FF;
            if constexpr (mode_exponent(D, K, f, d) == 0) {
                constexpr int fs = face_slot(D, K, f, d);
                R Am, Bm, Ap, Bp;
                modal_trace<D, K, R, d, 0, f>(uP, Am, Bm);
                modal_trace<D, K, R, d, 1, f>(uP, Ap, Bp);
                const R hmA = halo ? lane_value(hv, 2 * fs) : R(0), hmB = halo ? lane_value(hv, 2 * fs + 1) : R(0);
                const R hpA = halo ? lane_value(hv, 32 + 2 * fs) : R(0), hpB = halo ? lane_value(hv, 33 + 2 * fs) : R(0);
                // '-' neighbour (lane - 1): its '+' traces; '+' neighbour (lane + 1): its '-' traces
                R AnM = wave_shift<0x138>(hmA, Ap), BnM = wave_shift<0x138>(hmB, Bp);
                R AnP = wave_shift<0x130>(hpA, Am), BnP = wave_shift<0x130>(hpB, Bm);
                const R CVm = cm.c1 * Bm + cm.c2 * BnM + cm.c3 * Am + cm.c4 * AnM;
                const R CDm = cm.c1 * Am + cm.c5 * AnM;
                const R CVp = cp.c1 * Bp + cp.c2 * BnP + cp.c3 * Ap + cp.c4 * AnP;
                const R CDp = cp.c1 * Ap + cp.c5 * AnP;
                modal_scatter<D, K, R, d, 0, f>(CVm, CDm, y);
                modal_scatter<D, K, R, d, 1, f>(CVp, CDp, y);
            }
        });
    }
    // 4. DZ faces (3D): the neighbours' modes from memory, one side at a time
    if constexpr (D == 3) {
        constexpr int d = 2;
        const R ihd = ih[d];
        static_for<2>([&](auto SS) {
            constexpr int s = SS;
            const auto g = side_geom<D, R, d, s>(a, idx, nn, cjP, ihd, pen);
            R un[Np];
            const R nuQ = a.nu[g.cq];
            if constexpr (PRE) {
#pragma unroll
                for (int n = 0; n < Np; ++n) un[n] = zq[s][n];
            } else {
                const R* uq = a.u + (size_t)g.cq * Np;
#pragma unroll
                for (int n = 0; n < Np; ++n) un[n] = uq[n];
            }
            const SideCoef<R> cf = side_coef<R, s>(g.bnd, g.skip, R(2) * ihd, g.gQ, g.rq, nuP, nuQ, g.mu);
            face_side_traces<D, K, R, d, s>(uP, cf, [&](auto F, auto, R& An, R& Bn) {
                modal_trace<D, K, R, d, 1 - s, decltype(F)::value>(un, An, Bn);
            }, y);
        });
    }
    // 5. DT faces: the adjacent waves' traces from LDS, the workgroup's edge rows from memory;
    // the '-' side reads wave w - 1's '+' traces, then every wave replaces its '+' traces by
    // its '-' traces, which the '+' side of wave w + 1 reads
    {
        constexpr int d = DT;
        const R ihd = ih[d];
        auto dt_side = [&](auto SS) {
            constexpr int s = SS;
            const auto g = side_geom<D, R, d, s>(a, idx, nn, cjP, ihd, pen);
            const R nuQ = a.nu[g.cq];
            const SideCoef<R> cf = side_coef<R, s>(g.bnd, g.skip, R(2) * ihd, g.gQ, g.rq, nuP, nuQ, g.mu);
            const int wn = s ? w + 1 : w - 1;
            if (g.bnd || (wn >= 0 && wn < W)) {   // wave-uniform
                const int wr = g.bnd ? w : wn;
                face_side_traces<D, K, R, d, s>(uP, cf, [&](auto, auto FS, R& An, R& Bn) {
                    constexpr int fs = decltype(FS)::value;
                    An = tr[wr][2 * fs][lane];   // on a boundary side: finite, multiplied by zero
                    Bn = tr[wr][2 * fs + 1][lane];
                }, y);
            } else {
                R un[Np];
                if constexpr (PRE) {
#pragma unroll
                    for (int n = 0; n < Np; ++n) un[n] = tq[n];
                } else {
                    const R* uq = a.u + (size_t)g.cq * Np;
#pragma unroll
                    for (int n = 0; n < Np; ++n) un[n] = uq[n];
                }
                face_side_traces<D, K, R, d, s>(uP, cf, [&](auto F, auto, R& An, R& Bn) {
                    modal_trace<D, K, R, d, 1 - s, decltype(F)::value>(un, An, Bn);
                }, y);
            }
        };
        __syncthreads();   // every wave's '+' traces are in LDS
        dt_side(std::integral_constant<int, 0>{});
        __syncthreads();   // ... and have been read
        put_traces(std::integral_constant<int, 0>{});
        __syncthreads();
        dt_side(std::integral_constant<int, 1>{});
    }
    // 6. results: the wave's row of cells is contiguous in memory; after the barrier no wave
    // reads the trace buffers any more, so each wave stages its results in its own
    __syncthreads();
    const int nvalid = r < nn[DT] ? (nn[DF] - x0 < 64 ? nn[DF] - x0 : 64) : 0;
    int idx0[3] = {idx[0], idx[1], idx[2]};
    idx0[DF] = x0;
    store_cells_coalesced<Np, R>(&tr[w][0][0], y, lane, (size_t)encode<D>(idx0, a.n0, a.n1), nvalid, a);
}

// Assembly: thread = (cell r, trial mode n).  Column n of the diagonal block
// is the apply with u = e_n on r only; column n of the block coupling r to
// its neighbour across face f is the face-f term with u_P = 0, u_Q = e_n.
// Blocks are stored column-major, so each thread writes contiguous columns.
// Follows the split of LECVolumeQuadrature2 / LECEdgeQuadrature2
// (F/LECVolumeQuadrature2.cs:643-1140, F/LECEdgeQuadrature2.cs:1370-1560).
template <int D, int K>
__global__ __launch_bounds__(256) void sip_assemble_kernel(SipArgs<double> a, const int64_t* __restrict__ rowptr,
                                                           const int* __restrict__ col, double* __restrict__ vals) {
    constexpr int Np = binom(K + D, D);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.J * Np) return;
    const int c = (int)(t / Np), n = (int)(t % Np);
    int idx[3];
    decode<D>(c, a.n0, a.n1, idx);
    const int nn[3] = {a.n0, a.n1, a.n2};
    double h[3];
    cell_geom<D>(a, idx, h);
    double e[Np], z[Np], y[Np];
#pragma unroll
    for (int m = 0; m < Np; ++m) { e[m] = (m == n) ? 1.0 : 0.0; z[m] = 0; y[m] = 0; }
    double detJ = 1;
    for (int d = 0; d < D; ++d) detJ *= 0.5 * h[d];
    const double nuP = a.nu[c];
    double s2[D];
    for (int d = 0; d < D; ++d) s2[d] = 4.0 / (h[d] * h[d]);
    volume_term<D, K, double>(a, e, s2, nuP, y);
    const double cjP = cell_cj<D>(a, idx, h);
    const double sdPinv = 1.0 / sqrt(detJ);
    const int64_t r0 = rowptr[c], r1 = rowptr[c + 1];
    for (int f = 0; f < 2 * D; ++f) {
        const int d = f >> 1, s = f & 1;
        const bool isBnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
        const double dA = detJ / (0.5 * h[d]);
        const double gP = 2.0 / h[d];
        if (isBnd) {
            if ((a.bc_neumann >> f) & 1) continue;
            face_term<D, K, double, true, false>(a, f, e, z, sdPinv, 0.0, gP, 0.0, dA, nuP, nuP, a.pen * cjP, true, y);
        } else {
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] += s ? 1 : -1;
            const int cq = encode<D>(nidx, a.n0, a.n1);
            double hq[3] = {h[0], h[1], h[2]};
            hq[d] = a.width[d * a.nmax + nidx[d]];
            const double detQ = detJ / h[d] * hq[d];
            const double cjQ = cell_cj<D>(a, nidx, hq);
            const double mu = a.pen * fmax(cjP, cjQ);
            const double nuQ = a.nu[cq];
            // self part into the diagonal block column
            face_term<D, K, double, true, false>(a, f, e, z, sdPinv, 1.0 / sqrt(detQ), gP, 2.0 / hq[d], dA, nuP, nuQ, mu,
                                                 false, y);
            // coupling block (c, cq), column n
            double yo[Np];
#pragma unroll
            for (int m = 0; m < Np; ++m) yo[m] = 0;
            face_term<D, K, double, false, true>(a, f, z, e, sdPinv, 1.0 / sqrt(detQ), gP, 2.0 / hq[d], dA, nuP, nuQ, mu,
                                                 false, yo);
            int64_t b = -1;
            for (int64_t p = r0; p < r1; ++p) if (col[p] == cq) b = p;
            if (b >= 0) {
                double* dst = vals + (size_t)b * Np * Np + (size_t)n * Np;
#pragma unroll
                for (int m = 0; m < Np; ++m) dst[m] = yo[m];
            }
        }
    }
    int64_t b = -1;
    for (int64_t p = r0; p < r1; ++p) if (col[p] == c) b = p;
    if (b >= 0) {
        double* dst = vals + (size_t)b * Np * Np + (size_t)n * Np;
#pragma unroll
        for (int m = 0; m < Np; ++m) dst[m] = y[m];
    }
}

template <typename R>
static SipArgs<R> make_args(bh_op* op, const void* u, void* y, double alpha, double beta, bool add_affine) {
    SipArgs<R> a{};
    bh_grid* g = op->grid;
    a.J = g->J;
    a.n0 = g->n[0]; a.n1 = g->n[1]; a.n2 = g->n[2];
    a.nmax = g->nmax;
    a.width = g->d_width.as<double>();
    const bool f32 = sizeof(R) == 4;
    a.nu = (const R*)(f32 ? op->nu32.p : op->nu.p);
    const char* tb = (const char*)(f32 ? op->tab32.p : op->tab64.p);
    const size_t es = sizeof(R);
    a.G = (const R*)(tb + op->off_G * es);
    a.T = (const R*)(tb + op->off_T * es);
    a.Nn = (const R*)(tb + op->off_N * es);
    a.vw = (const R*)(tb + op->off_vw * es);
    a.fw = (const R*)(tb + op->off_fw * es);
    a.Kv = (const R*)(tb + op->off_K * es);
    a.pen = op->penalty_base * op->penalty_deg;
    a.bc_neumann = 0;
    for (int f = 0; f < 2 * op->D; ++f) if (op->bc[f] == 1) a.bc_neumann |= 1 << f;
    a.u = (const R*)u;
    a.y = (R*)y;
    a.affine = (add_affine && op->has_affine) ? op->affine.as<double>() : nullptr;
    a.alpha = alpha; a.beta = beta;
    return a;
}

// Vector-kernel forms: 1 = face rule node by node, 3 = modal faces, 4 = tiled modal
// faces (tile_w waves per workgroup).
template <typename R, int D, int K, int W>
static void launch_tile(bh_op* op, const SipArgs<R>& a) {
    const int DF = D == 2 ? 1 : 0, DT = D == 2 ? 0 : 1;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int nseg = (nn[DF] + 63) / 64, nrb = (nn[DT] + W - 1) / W, nz = D == 3 ? nn[2] : 1;
    const int64_t nt = (int64_t)nseg * nrb * nz;
    BH_CHECK(nt < (int64_t)1 << 30, "sip_apply: grid too large for the tiled kernel");
    const int ntiles = (int)nt, per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL((sip_apply_tile_kernel<D, K, R, W>), dim3(8 * per_xcd), dim3(64 * W), 0, op->ctx->stream, a, nseg,
                       nrb, ntiles, per_xcd);
}

template <typename R, int D, int K>
static void launch_apply(bh_op* op, const SipArgs<R>& a, int form, int tile_w) {
    dim3 blk(256), grd((a.J + 255) / 256);
    if (form == 4) {
        if (tile_w == 2) launch_tile<R, D, K, 2>(op, a);
        else if (tile_w == 8) launch_tile<R, D, K, 8>(op, a);
        else launch_tile<R, D, K, 4>(op, a);
    } else if (form == 3) {
        hipLaunchKernelGGL((sip_apply_kernel<D, K, R, true>), grd, blk, 0, op->ctx->stream, a);
    } else {
        hipLaunchKernelGGL((sip_apply_kernel<D, K, R, false>), grd, blk, 0, op->ctx->stream, a);
    }
    BH_HIP(hipGetLastError());
}

template <typename R>
static void dispatch_apply(bh_op* op, const SipArgs<R>& a, int form, int tile_w) {
    const int D = op->D, K = op->k;
#define BH_CASE(DD, KK) if (D == DD && K == KK) return launch_apply<R, DD, KK>(op, a, form, tile_w);
    BH_CASE(2, 1) BH_CASE(2, 2) BH_CASE(2, 3) BH_CASE(2, 4)
    BH_CASE(3, 1) BH_CASE(3, 2) BH_CASE(3, 3)
#undef BH_CASE
    throw Error(-3, "sip_apply: unsupported (D, k) = (" + std::to_string(D) + ", " + std::to_string(K) + ")");
}

namespace {
struct ApplyTimer {
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit ApplyTimer(hipStream_t s) : st(s) {
        BH_HIP(hipEventCreate(&e0));
        BH_HIP(hipEventCreate(&e1));
    }
    ~ApplyTimer() {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    float operator()(const std::function<void()>& f) {
        f();
        BH_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 3; ++r) f();
        BH_HIP(hipEventRecord(e1, st));
        BH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        BH_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms / 3;
    }
};
}  // namespace

// Times the vector-kernel forms of precision R (modal, tiled modal with 2, 4 and 8
// waves per workgroup, and node by node when with_node) on scratch vectors of this
// operator; returns the fastest form and its time.
template <typename R>
static float tune_vector_forms(bh_op* op, bool with_node, int& form, int& tile_w) {
    DevBuf us(op->N * sizeof(R)), ys(op->N * sizeof(R));
    if (sizeof(R) == 8) fill(op->ctx, us.as<double>(), 1.0, op->N);
    else fill32(op->ctx, us.as<float>(), 1.0f, op->N);
    ApplyTimer time_it(op->ctx->stream);
    const SipArgs<R> a = make_args<R>(op, us.p, ys.p, 1.0, 0.0, false);
    float best = time_it([&] { dispatch_apply<R>(op, a, 3, 0); });
    form = 3;
    tile_w = 0;
    for (int w : {2, 4, 8}) {
        const float t = time_it([&] { dispatch_apply<R>(op, a, 4, w); });
        if (t < best) { best = t; form = 4; tile_w = w; }
    }
    if (with_node) {
        const float t = time_it([&] { dispatch_apply<R>(op, a, 1, 0); });
        if (t < best) { best = t; form = 1; tile_w = 0; }
    }
    return best;
}

static void tune_fp64_path_timed(bh_op* op);

// Tuning decisions are kept per (device, dimension, degree, grid shape, precision) for
// the process: an operator of a shape already tuned takes the earlier choice without
// timing again, so every operator of one shape runs the same kernel and equal inputs
// give bitwise equal outputs across operator instances (the forms agree only up to
// rounding).  Re-arming the tuner of an operator (path 0) forgets its shape's entry.
namespace {
struct TuneRec {
    int path = 0, tile_w = 0, waves = 0, wg_per_xcd = 0;
    float ms_vector = 0.f, ms_mfma = 0.f;
};
using TuneKey = std::array<int, 7>;
std::mutex tune_mu;
std::map<TuneKey, TuneRec> tune_cache;
TuneKey tune_key(const bh_op* op, int prec) {
    const bh_grid* g = op->grid;
    return {op->ctx->device, op->D, op->k, g->n[0], g->n[1], g->n[2], prec};
}
}  // namespace

void sip_forget_tuning(bh_op* op, int prec) {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache.erase(tune_key(op, prec));
}

// Autotuner of the FP64 apply: times the vector kernel (face terms node by node,
// modal, tiled modal) and the FP64-MFMA kernel (waves per workgroup x workgroups
// per XCD) on a scratch vector of this operator and keeps the fastest; all
// compute the same integrals, up to rounding.
static void tune_fp64_path(bh_op* op) {
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        auto it = tune_cache.find(tune_key(op, BH_FP64));
        if (it != tune_cache.end()) {
            const TuneRec& t = it->second;
            op->f64_path = t.path;
            op->f64_tile_w = t.tile_w;
            op->f64_waves = t.waves;
            op->f64_wg_per_xcd = t.wg_per_xcd;
            op->f64_ms_vector = t.ms_vector;
            op->f64_ms_mfma = t.ms_mfma;
            return;
        }
    }
    tune_fp64_path_timed(op);
    TuneRec t;
    t.path = op->f64_path;
    t.tile_w = op->f64_tile_w;
    t.waves = op->f64_waves;
    t.wg_per_xcd = op->f64_wg_per_xcd;
    t.ms_vector = op->f64_ms_vector;
    t.ms_mfma = op->f64_ms_mfma;
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache[tune_key(op, BH_FP64)] = t;
}

static void tune_fp64_path_timed(bh_op* op) {
    int form = 3, tw = 0;
    const float tv = tune_vector_forms<double>(op, true, form, tw);
    op->f64_path = form;
    op->f64_tile_w = tw;
    op->f64_ms_vector = tv;
    if (!sip_apply_mfma64_supported(op->D, op->k, op->N)) return;
    DevBuf us(op->N * sizeof(double)), ys(op->N * sizeof(double));
    fill(op->ctx, us.as<double>(), 1.0, op->N);
    ApplyTimer time_it(op->ctx->stream);
    float best = 1e30f;
    int bw = 8, bx = 64;
    for (int w : {4, 8})
        for (int px : {32, 64, 128}) {
            const float t = time_it([&] { sip_apply_mfma64(op, 1.0, us.as<double>(), 0.0, ys.as<double>(), nullptr, w, px); });
            if (t < best) { best = t; bw = w; bx = px; }
        }
    op->f64_ms_mfma = best;
    op->f64_waves = bw;
    op->f64_wg_per_xcd = bx;
    if (best < tv) op->f64_path = 2;
}

// The FP32 apply (multigrid levels) is tuned on its own, over the modal forms.
static void tune_fp32_path(bh_op* op) {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache.find(tune_key(op, BH_FP32));
    if (it != tune_cache.end()) {
        op->f32_path = it->second.path;
        op->f32_tile_w = it->second.tile_w;
        op->f32_ms = it->second.ms_vector;
        return;
    }
    int form = 3, tw = 0;
    op->f32_ms = tune_vector_forms<float>(op, false, form, tw);
    op->f32_path = form;
    op->f32_tile_w = tw;
    TuneRec t;
    t.path = form;
    t.tile_w = tw;
    t.ms_vector = op->f32_ms;
    tune_cache[tune_key(op, BH_FP32)] = t;
}

void sip_apply(bh_op* op, double alpha, const void* u, double beta, void* y, bool add_affine, int prec, const void* b0) {
    if (add_affine && op->has_affine && !op->affine_valid) op->rebuild_affine();
    if (prec == BH_FP64) {
        if (op->f64_path == 0) tune_fp64_path(op);
        if (op->f64_path == 2) {
            const double* aff = (add_affine && op->has_affine) ? op->affine.as<double>() : nullptr;
            sip_apply_mfma64(op, alpha, (const double*)u, beta, (double*)y, aff, op->f64_waves, op->f64_wg_per_xcd,
                             (const double*)b0);
            return;
        }
        SipArgs<double> a = make_args<double>(op, u, y, alpha, beta, add_affine);
        a.b0 = (const double*)b0;
        dispatch_apply<double>(op, a, op->f64_path, op->f64_tile_w);
    } else if (prec == BH_FP32) {
        if (add_affine && op->has_affine) throw Error(-3, "FP32 evaluate supports the linear part only");
        if (op->f32_path == 0) tune_fp32_path(op);
        SipArgs<float> a = make_args<float>(op, u, y, alpha, beta, false);
        a.b0 = (const float*)b0;
        dispatch_apply<float>(op, a, op->f32_path, op->f32_tile_w);
    } else if (prec == BH_FP16_MFMA) {
        sip_apply_mfma(op, alpha, (const float*)u, beta, (float*)y, add_affine, (const float*)b0);
    } else {
        throw Error(-3, "sip_apply: precision not supported by this entry point");
    }
}

template <int D, int K>
static void launch_asm(bh_op* op, bh_bsr* M) {
    SipArgs<double> a = make_args<double>(op, nullptr, nullptr, 1.0, 0.0, false);
    constexpr int Np = binom(K + D, D);
    int64_t nt = (int64_t)a.J * Np;
    dim3 blk(256), grd((unsigned)((nt + 255) / 256));
    hipLaunchKernelGGL((sip_assemble_kernel<D, K>), grd, blk, 0, op->ctx->stream, a, M->rowptr.as<int64_t>(),
                       M->col.as<int>(), M->vals.as<double>());
    BH_HIP(hipGetLastError());
}

void sip_assemble(bh_op* op, bh_bsr* M) {
    const int D = op->D, K = op->k;
#define BH_CASE(DD, KK) if (D == DD && K == KK) return launch_asm<DD, KK>(op, M);
    BH_CASE(2, 1) BH_CASE(2, 2) BH_CASE(2, 3) BH_CASE(2, 4)
    BH_CASE(3, 1) BH_CASE(3, 2) BH_CASE(3, 3)
#undef BH_CASE
    throw Error(-3, "sip_assemble: unsupported (D, k)");
}

}  // namespace bh
