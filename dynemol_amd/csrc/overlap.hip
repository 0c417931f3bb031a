// STO overlap S and Wolfsberg-Helmholz Hueckel h, built on the device.
//
// Follows the reference's two-centre scheme (overlap_D.f:144-311): for every
// atom pair (ia > ib) within the cutoff, lined-up-frame overlaps
// <n l m | n' l' m> from the Fernandez-Rico basic integrals and the L, L', M
// recursions (SOLAP, overlap_D.f:658-962), STO normalisation (GET_ANOR,
// :466-517), then rotation to the molecular frame with real-harmonic rotation
// matrices (ROTAR/DLMN, :983-1191).  h_ij = X_ij * S_ij (hamiltonians.f:33-63).
//
// Work decomposition: one thread per (atom pair, shell pair).  A shell is the
// set of 2l+1 orbitals sharing (n, l, zeta, coef) on one atom; the lined-up
// integrals depend only on the shell pair, so each thread evaluates them once
// and writes the (2la+1) x (2lb+1) block of S and h (and the mirrored block).
// The reference evaluates SOLAP once per orbital pair instead.
#include "common.hpp"
#include "kernels.hpp"

namespace dh {

// Bounds for the compact device tables: n <= 5, l <= 2 (s, p, d shells).
constexpr int SQ_NMAX = 5;
constexpr int SQ_LMAX = 2;
constexpr int SQ_D = SQ_NMAX + 3 * SQ_LMAX + 1;  // smat index range 0..SQ_D-1 (nmax <= n+l+2l' = 11)
constexpr int SQ_HA = SQ_D + 2;                   // 1F1 'a' index range
constexpr int SQ_HG = 2 * SQ_D + 3;               // 1F1 'g' index range (ngmax <= 24)

__device__ __forceinline__ double rre(int i) { return (double)i; }
__device__ __forceinline__ double rreali(int i) { return i == 0 ? 0.0 : 1.0 / (double)i; }

__device__ double dfact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= (double)i;
    return f;
}

// Lined-up overlaps of unnormalised STOs, sol[0..min(la,lb)].
__device__ void solap_dev(int na, int la, double za, int nb, int lb, double zb, double R, double* sol) {
    const double pi4 = 12.566370614359172953850573533118012;
    double h1f1[SQ_HA * SQ_HG];
    double smat[SQ_D * SQ_D], smataux[SQ_D * SQ_D], smatm[SQ_D * SQ_D];
#define H1(a, g) h1f1[(a) * SQ_HG + (g)]
#define SM(n, np) smat[(n) * SQ_D + (np)]
#define SA(n, np) smataux[(n) * SQ_D + (np)]
#define SMM(n, np) smatm[(n) * SQ_D + (np)]
    // h1f1 needs no clearing: level ngmax is filled for a = 1..namax, level j
    // for a = 1..min(namax, j - 1) from level j + 1, and every later read
    // (j + 1, k + 2) has j + 1 <= min(namax, k + 1).  Nor do smat, smataux and
    // smatm: the basic integrals fill n = nmin..nmax, n' = n'min..n'max - (n -
    // nmin), and every read of the L, L' and M recursions that follow lands on
    // an entry written before it (checked over every n <= 5, l <= 2 shell
    // pair in both zeta orders, all indices within the 12 x 12 tables).
    int nmin, nmax, npmin, npmax;
    double z, zp;
    const bool a_first = za >= zb;
    if (a_first) {
        nmin = na - la; nmax = na + la + 2 * lb;
        npmin = nb - lb; npmax = nb + lb + 2 * la;
        z = za; zp = zb;
    } else {
        nmin = nb - lb; nmax = nb + lb + 2 * la;
        npmin = na - la; npmax = na + la + 2 * lb;
        z = zb; zp = za;
    }
    const double expzR = exp(-z * R);
    const int ngmax = nmax + npmax + 2;
    const int namax = npmax + 1;
    const double arg = (z - zp) * R;
    const int iz = (int)z;
    int ia = (ngmax - iz) / 2;
    if (ia < 1) ia = 1;
    if (ia >= namax) ia = namax - 1;

    // 1F1[ia, ngmax, arg] and 1F1[ia+1, ngmax, arg] by their series
    double auxa = arg * rre(ia) * rreali(ngmax), f11a = 1.0;
    double auxb = arg * rre(ia + 1) * rreali(ngmax), f11b = 1.0;
    for (int i = 1; i <= 1000; ++i) {
        f11a += auxa;
        f11b += auxb;
        auxa = auxa * rre(ia + i) * arg * rreali(ngmax + i) * rreali(i + 1);
        auxb = auxb * rre(ia + i + 1) * arg * rreali(ngmax + i) * rreali(i + 1);
        if (auxb < 1.0e-35 * f11b) break;
    }
    H1(ia, ngmax) = f11a;
    H1(ia + 1, ngmax) = f11b;
    double aux = arg - rre(ngmax);
    for (int i = ia + 1; i <= namax - 1; ++i)
        H1(i + 1, ngmax) = ((aux + rre(2 * i)) * H1(i, ngmax) + rre(ngmax - i) * H1(i - 1, ngmax)) * rreali(i);
    for (int i = ia; i >= 2; --i)
        H1(i - 1, ngmax) = (rre(i) * H1(i + 1, ngmax) - (aux + rre(2 * i)) * H1(i, ngmax)) * rreali(ngmax - i);
    for (int j = ngmax - 1; j >= 2; --j) {
        H1(1, j) = (H1(2, j + 1) + rre(j - 1) * H1(1, j + 1)) * rreali(j);
        aux = arg * rreali(j);
        const int top = min(namax - 1, j - 2);
        for (int i = 1; i <= top; ++i) H1(i + 1, j) = aux * H1(i + 1, j + 1) + H1(i, j);
    }

    // basic integrals <N00|N'00>
    {
        const bool mcopy = min(la, lb) > 0;
        const double zzpR = (z + zp) * R;
        const double zzpi = 1.0 / (z + zp);
        double zzpin = 1.0;
        for (int e = 0; e < nmin + npmin + 1; ++e) zzpin *= zzpi;
        for (int n = nmin; n <= nmax; ++n) {
            double zzpinp = zzpin;
            for (int np = npmin; np <= npmax - n + nmin; ++np) {
                double auxk = dfact(n + np) * zzpinp;
                double sumk = 0.0;
                for (int k = 0; k <= n + np; ++k) {
                    double sumj = 0.0;
                    for (int j = max(0, k - n); j <= min(k, np); ++j)
                        sumj += H1(j + 1, k + 2) / (dfact(n + j - k) * dfact(np - j));
                    sumk += sumj * auxk;
                    auxk = auxk * zzpR * rreali(n + np - k) * rreali(k + 2);
                }
                const double v = pi4 * expzR * dfact(n) * dfact(np) * sumk;
                // smataux needs no copy (every read of it sees a recursion
                // write), smatm only when an M step follows (min(la, lb) > 0)
                if (a_first) {
                    SM(n, np) = v;
                    if (mcopy) SMM(n, np) = v;
                } else {
                    SM(np, n) = v;
                    if (mcopy) SMM(np, n) = v;
                }
                zzpinp *= zzpi;
            }
            zzpin *= zzpi;
        }
    }

    nmin = na - la; nmax = na + la + 2 * lb;
    npmin = nb - lb; npmax = nb + lb + 2 * la;
    const double Rinv = 1.0 / R, R2 = R * R, R4 = R2 * R2;
    double updltm0i = 0.5;
    const int mtop = min(la, lb);
    for (int m = 0; m <= mtop; ++m) {
        if (la > m) {
            for (int np = npmin; np <= npmax - 2; ++np) {
                double saux = SM(nmin, np);
                for (int n = nmin + 1; n <= nmax - 1 - np + npmin; ++n) {
                    const double sbux = rre(2 * m + 1) * 0.5 * Rinv * (SM(n + 1, np) + R2 * saux - SM(n - 1, np + 2));
                    saux = SM(n, np);
                    SA(n, np) = saux;
                    SM(n, np) = sbux;
                }
            }
            for (int ll = m + 1; ll <= la - 1; ++ll)
                for (int np = npmin; np <= npmax - 2 * (ll + 1 - m); ++np) {
                    double saux = SM(nmin + ll - m, np);
                    for (int n = nmin + 1 + ll - m; n <= nmax - 1 - np + npmin - ll + m; ++n) {
                        const double sbux =
                            rre(2 * ll + 1) * 0.5 * rreali(ll - m + 1) * Rinv * (SM(n + 1, np) + R2 * saux - SM(n - 1, np + 2)) -
                            rre(ll + m) * rreali(ll - m + 1) * SA(n, np);
                        saux = SM(n, np);
                        SA(n, np) = saux;
                        SM(n, np) = sbux;
                    }
                }
        }
        if (lb > m) {
            for (int n = nmin + la - m; n <= nmax - la + m - 2; ++n) {
                double saux = SM(n, npmin);
                for (int np = npmin + 1; np <= npmax - 2 * (la - m) - n + nmin + la - m - 1; ++np) {
                    const double sbux = -rre(2 * m + 1) * 0.5 * Rinv * (SM(n, np + 1) + R2 * saux - SM(n + 2, np - 1));
                    saux = SM(n, np);
                    SA(n, np) = saux;
                    SM(n, np) = sbux;
                }
            }
            for (int ll = m + 1; ll <= lb - 1; ++ll)
                for (int n = nmin + la - m; n <= nmax - la + m - 2 * (ll + 1 - m); ++n) {
                    double saux = SM(n, npmin + ll - m);
                    for (int np = npmin + 1 + ll - m; np <= npmax - 2 * (la - m) - n + nmin + la - m - ll + m - 1; ++np) {
                        const double sbux =
                            -rre(2 * ll + 1) * 0.5 * rreali(ll - m + 1) * Rinv * (SM(n, np + 1) + R2 * saux - SM(n + 2, np - 1)) -
                            rre(ll + m) * rreali(ll - m + 1) * SA(n, np);
                        saux = SM(n, np);
                        SA(n, np) = saux;
                        SM(n, np) = sbux;
                    }
                }
        }
        sol[m] = SM(na, nb);
        if (m == mtop) break;
        nmin += 1; nmax -= 3; npmin += 1; npmax -= 3;
        aux = updltm0i * 0.25 * rre(2 * m + 1) * rre(2 * m + 1) * Rinv * Rinv;
        for (int n = nmin; n <= nmax; ++n)
            for (int np = npmin; np <= npmax - n + nmin; ++np)
                SA(n, np) = aux * (2.0 * (SMM(n + 1, np + 1) + R2 * (SMM(n + 1, np - 1) + SMM(n - 1, np + 1))) -
                                   SMM(n + 3, np - 1) - SMM(n - 1, np + 3) - R4 * SMM(n - 1, np - 1));
        for (int n = nmin; n <= nmax; ++n)
            for (int np = npmin; np <= npmax - n + nmin; ++np) {
                SM(n, np) = SA(n, np);
                SMM(n, np) = SA(n, np);
            }
        updltm0i = 1.0;
    }
#undef H1
#undef SM
#undef SA
#undef SMM
}

// Real-harmonic rotation matrices up to l = 2 for the A->B axis (gamma = 0).
// rl[(mp+2) + 5*((m+2) + 5*l)], l = 0..2.
#define RX(mp, m, l) (((mp) + 2) + 5 * (((m) + 2) + 5 * (l)))
__device__ void rotation_dev(double xab, double yab, double zab, int lmax, double* rl, double* Rab_out) {
    const double root2 = 1.414213562373095;
    double dl[75];
    for (int i = 0; i < 75; ++i) {
        dl[i] = 0.0;
        rl[i] = 0.0;
    }
    const double xy = sqrt(xab * xab + yab * yab);
    double sinal, cosal;
    if (xy > 1.0e-10) {
        sinal = yab / xy;
        cosal = xab / xy;
    } else {
        sinal = 0.0;
        cosal = 1.0;
    }
    const double Rab = sqrt(xab * xab + yab * yab + zab * zab);
    *Rab_out = Rab;
    double sinbet = xy / Rab, cosbet = zab / Rab;
    dl[RX(0, 0, 0)] = 1.0;
    rl[RX(0, 0, 0)] = 1.0;
    if (lmax == 0) return;
    dl[RX(1, 1, 1)] = (1.0 + cosbet) * 0.5;
    dl[RX(1, 0, 1)] = -sinbet / root2;
    dl[RX(1, -1, 1)] = (1.0 - cosbet) * 0.5;
    dl[RX(0, 1, 1)] = -dl[RX(1, 0, 1)];
    dl[RX(0, 0, 1)] = dl[RX(1, 1, 1)] - dl[RX(1, -1, 1)];
    dl[RX(0, -1, 1)] = dl[RX(1, 0, 1)];
    dl[RX(-1, 1, 1)] = dl[RX(1, -1, 1)];
    dl[RX(-1, 0, 1)] = dl[RX(0, 1, 1)];
    dl[RX(-1, -1, 1)] = dl[RX(1, 1, 1)];
    // gamma = 0: cos(ga) = 1, sin(ga) = 0
    rl[RX(0, 0, 1)] = dl[RX(0, 0, 1)];
    rl[RX(1, 0, 1)] = root2 * dl[RX(0, 1, 1)] * cosal;
    rl[RX(-1, 0, 1)] = root2 * dl[RX(0, 1, 1)] * sinal;
    rl[RX(0, 1, 1)] = root2 * dl[RX(1, 0, 1)];
    rl[RX(0, -1, 1)] = 0.0;
    rl[RX(1, 1, 1)] = dl[RX(1, 1, 1)] * cosal - dl[RX(1, -1, 1)] * cosal;
    rl[RX(1, -1, 1)] = -dl[RX(1, 1, 1)] * sinal - dl[RX(1, -1, 1)] * sinal;
    rl[RX(-1, 1, 1)] = dl[RX(1, 1, 1)] * sinal - dl[RX(1, -1, 1)] * sinal;
    rl[RX(-1, -1, 1)] = dl[RX(1, 1, 1)] * cosal + dl[RX(1, -1, 1)] * cosal;
    if (lmax < 2) return;
    double tgbet2;
    if (fabs(sinbet) < 1.0e-10) {
        tgbet2 = 0.0;
        sinbet = 0.0;
        cosbet = copysign(1.0, cosbet);
    } else {
        tgbet2 = (1.0 - cosbet) / sinbet;
    }
    // DLMN for l = 2
    const int l = 2;
    int iinf = 1 - l, isup = -iinf;
    auto root = [](int i) { return sqrt((double)i); };
    auto rooti = [](int i) { return 1.0 / sqrt((double)i); };
    dl[RX(l, l, l)] = dl[RX(isup, isup, l - 1)] * (1.0 + cosbet) * 0.5;
    dl[RX(l, -l, l)] = dl[RX(isup, -isup, l - 1)] * (1.0 - cosbet) * 0.5;
    for (int m = isup; m >= iinf; --m) dl[RX(l, m, l)] = -tgbet2 * root(l + m + 1) * rooti(l - m) * dl[RX(l, m + 1, l)];
    const double al = l, al1 = al - 1.0, tal1 = al + al1, ali = 1.0 / al1, cosaux = cosbet * al * al1;
    for (int mp = l - 1; mp >= 0; --mp) {
        const int laux = l + mp, lbux = l - mp;
        const double a2 = rooti(laux) * rooti(lbux) * ali;
        const double cux = root(laux - 1) * root(lbux - 1) * al;
        for (int m = isup; m >= iinf; --m) {
            const int lauz = l + m, lbuz = l - m;
            const double factor = a2 * rooti(lauz) * rooti(lbuz);
            double term = tal1 * (cosaux - (double)m * (double)mp) * dl[RX(mp, m, l - 1)];
            if (lbuz != 1 && lbux != 1) term -= dl[RX(mp, m, l - 2)] * cux * root(lauz - 1) * root(lbuz - 1);
            dl[RX(mp, m, l)] = factor * term;
        }
        iinf += 1;
        isup -= 1;
    }
    double sign = 1.0;
    iinf = -l;
    isup = l - 1;
    for (int m = l; m >= 1; --m) {
        for (int mp = iinf; mp <= isup; ++mp) {
            dl[RX(mp, m, l)] = sign * dl[RX(m, mp, l)];
            sign = -sign;
        }
        iinf += 1;
        isup -= 1;
    }
    iinf = -l;
    isup = iinf;
    for (int m = l - 1; m >= -l; --m) {
        sign = -1.0;
        for (int mp = isup; mp >= iinf; --mp) {
            dl[RX(mp, m, l)] = sign * dl[RX(-mp, -m, l)];
            sign = -sign;
        }
        isup += 1;
    }
    rl[RX(0, 0, l)] = dl[RX(0, 0, l)];
    double cosmal = cosal, sinmal = sinal;
    sign = -1.0;
    for (int mp = 1; mp <= l; ++mp) {
        const double a0 = root2 * dl[RX(0, mp, l)];
        rl[RX(mp, 0, l)] = a0 * cosmal;
        rl[RX(-mp, 0, l)] = a0 * sinmal;
        for (int m = 1; m <= l; ++m) {
            const double a1 = root2 * dl[RX(m, 0, l)];
            rl[RX(0, m, l)] = a1;
            rl[RX(0, -m, l)] = 0.0;
            const double d1 = dl[RX(-mp, -m, l)];
            const double d2 = sign * dl[RX(mp, -m, l)];
            rl[RX(mp, m, l)] = d1 * cosmal + d2 * cosmal;
            rl[RX(mp, -m, l)] = -d1 * sinmal + d2 * sinmal;
            rl[RX(-mp, m, l)] = d1 * sinmal + d2 * sinmal;
            rl[RX(-mp, -m, l)] = d1 * cosmal - d2 * cosmal;
        }
        sign = -sign;
        const double t = cosmal * cosal - sinmal * sinal;
        sinmal = sinmal * cosal + cosmal * sinal;
        cosmal = t;
    }
}

// S = 0 with unit diagonal; h = 0 with h_ii = IP_i + V_shift_i.  Pad rows/cols
// in [N, ld) stay zero (the padded diagonal of S is set to one so that the
// padded Cholesky factor is the identity).
__global__ void init_SH_kernel(double* S, double* h, int N, int ld, const double* ip, const double* vshift) {
    const long total = (long)ld * ld;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % ld), j = (int)(t / ld);
        double s = 0.0, hv = 0.0;
        if (i == j) {
            s = 1.0;
            hv = i < N ? ip[i] + vshift[i] : 0.0;
        }
        S[t] = s;
        h[t] = hv;
    }
}

__global__ void build_SH_kernel(const BasisDev b, double* S, double* h, int ld, double cutoff, double s_drop) {
    // Shell-pair major, then pairs (p > r) of atoms in basis-signature order:
    // the 64 lanes of a wavefront share one shell pair and mostly one element
    // pair, so they follow the same SOLAP branches.
    const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long npair = (long)b.n_atoms * (b.n_atoms - 1) / 2;
    const int sp = (int)(tid / npair);
    const long q = tid - (long)sp * npair;
    if (sp >= MAX_SHELLS * MAX_SHELLS) return;
    // q -> (p > r): q = p*(p-1)/2 + r
    int p = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
    while ((long)p * (p - 1) / 2 > q) --p;
    while ((long)(p + 1) * p / 2 <= q) ++p;
    const int r = (int)(q - (long)p * (p - 1) / 2);
    const int ap = b.atom_order[p], ar = b.atom_order[r];
    // the reference's roles: A is the atom of higher index (overlap_D.f:144-311)
    const bool p_is_a = ap > ar;
    const int ia = p_is_a ? ap : ar, ib = p_is_a ? ar : ap;
    const int sa = p_is_a ? sp / MAX_SHELLS : sp % MAX_SHELLS;
    const int sb = p_is_a ? sp % MAX_SHELLS : sp / MAX_SHELLS;
    if (sa >= b.atom_nshell[ia] || sb >= b.atom_nshell[ib]) return;
    const double xab = b.coords[3 * ib + 0] - b.coords[3 * ia + 0];
    const double yab = b.coords[3 * ib + 1] - b.coords[3 * ia + 1];
    const double zab = b.coords[3 * ib + 2] - b.coords[3 * ia + 2];
    const double R = sqrt(xab * xab + yab * yab + zab * zab);
    if (R > cutoff) return;
    const int shA = b.atom_shell0[ia] + sa, shB = b.atom_shell0[ib] + sb;
    const int na = b.sh_n[shA], la = b.sh_l[shA], nb = b.sh_n[shB], lb = b.sh_l[shB];
    double rl[75];
    double Rab;
    rotation_dev(xab, yab, zab, max(la, lb), rl, &Rab);
    const double aux = 1.0 / (dfact(2 * na) * dfact(2 * nb));
    const int msup = min(la, lb);
    double solnorm[3] = {0.0, 0.0, 0.0};
    double solvec[3];
    for (int j = 0; j < b.sh_nzeta[shB]; ++j) {
        for (int i = 0; i < b.sh_nzeta[shA]; ++i) {
            const double expb = b.sh_zeta[2 * shB + j], expa = b.sh_zeta[2 * shA + i];
            solap_dev(na, la, expa, nb, lb, expb, Rab, solvec);
            // GET_ANOR
            double anor = pow(expa * 2.0, (double)(na * 2 + 1)) * pow(expb * 2.0, (double)(nb * 2 + 1)) *
                          (la * 2 + 1.0) * (lb * 2 + 1.0) * aux;
            anor = sqrt(anor) / (4.0 * 3.141592653589793238462643383279502884);
            double part[3];
            for (int m = 0; m <= msup - 1; ++m) {
                part[m] = solvec[m] * anor;
                anor = anor / sqrt((la + m + 1.0) * (la - m) * (lb + m + 1.0) * (lb - m));
                if (m == 0) anor = anor + anor;
            }
            part[msup] = solvec[msup] * anor;
            const double cc = b.sh_coef[2 * shA + i] * b.sh_coef[2 * shB + j];
            for (int k = 0; k <= msup; ++k) solnorm[k] += cc * part[k];
        }
    }
    const int a0 = b.sh_orb0[shA], b0 = b.sh_orb0[shB];
    for (int mb = -lb; mb <= lb; ++mb) {
        for (int ma = -la; ma <= la; ++ma) {
            double s = solnorm[0] * rl[RX(ma, 0, la)] * rl[RX(mb, 0, lb)];
            for (int k = 1; k <= msup; ++k)
                s += solnorm[k] * (rl[RX(ma, -k, la)] * rl[RX(mb, -k, lb)] + rl[RX(ma, k, la)] * rl[RX(mb, k, lb)]);
            if (fabs(s) < s_drop) s = 0.0;
            const int a = a0 + ma + la, bo = b0 + mb + lb;
            const double hv = huckel_x(b.ip[a], b.ip[bo], b.kwh[a], b.kwh[bo], b.vshift[a], b.vshift[bo]) * s;
            S[(long)a + (long)ld * bo] = s;
            S[(long)bo + (long)ld * a] = s;
            h[(long)a + (long)ld * bo] = hv;
            h[(long)bo + (long)ld * a] = hv;
        }
    }
}
#undef RX

int launch_build_SH(const BasisDev& b, int N, int ld, double* S, double* h, double cutoff, double s_drop,
                    hipStream_t st) {
    {
        const long total = (long)ld * ld;
        int blocks = (int)std::min<long>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(init_SH_kernel, dim3(blocks), dim3(256), 0, st, S, h, N, ld, b.ip, b.vshift);
    }
    const long npair = (long)b.n_atoms * (b.n_atoms - 1) / 2;
    if (npair > 0) {
        const long threads = npair * MAX_SHELLS * MAX_SHELLS;
        const int bs = 64;
        const long blocks = (threads + bs - 1) / bs;
        hipLaunchKernelGGL(build_SH_kernel, dim3((unsigned)blocks), dim3(bs), 0, st, b, S, h, ld, cutoff, s_drop);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace dh
