/*
 * CPU oracle for the Dynemol electron/hole wavepacket propagator path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dynemol_amd/, the HIP
 * library, the Fortran module) links, loads or calls this file.  It is used by
 * tests/ as the checker, by __graft_entry__.smoke() as the checker and by
 * bench.py's cpu_baseline leg as the timed CPU restatement.
 *
 * Every routine is a plain-C restatement of the reference algorithm, written
 * from the reference's behaviour, with the reference file:line it follows:
 *
 *   orc_util_init         overlap_D.f:1194-1231  (UTIL_OVERLAP tables)
 *   orc_solap             overlap_D.f:658-962    (SOLAP: lined-up STO overlaps)
 *   orc_anor/orc_anor_div overlap_D.f:466-517    (GET_ANOR, GET_ANOR_DIVISOR)
 *   orc_rotar             overlap_D.f:983-1191   (ROTAR + DLMN real-harmonic rotations)
 *   orc_overlap           overlap_D.f:144-311,520-583 (Build_Overlap_Matrix, RotationOverlap)
 *   orc_huckel            hamiltonians.f:33-63, ElHl_Chebyshev.f:296-323 (X_ij, Build_Huckel)
 *   orc_syinvert          GPU_Interface.cpp:936-949 + Matrix_math.f:126-198 (S^-1, symmetrised)
 *   orc_dgemm_nn          GPU_Interface.cpp:574-627 (xPU_dsymm: H' = S^-1 h)
 *   orc_propagation       Taylor.f:35-127 (Propagation), :132-219 (Convergence),
 *                         :224-239 (coefficient), :290-303 (isConverged)
 *   orc_populations       data_output.f:87-147,242-263 (Populations_mtx, pop_Slater)
 *   orc_qp_energy         ElHl_Chebyshev.f:329-371 (QuasiParticleEnergies)
 *   orc_huckel_x          hamiltonians.f:33-63, diabatic-Ehren.f:250-276 (Huckel_stuff: X_ij)
 *   orc_ehrenfest_kernel  diabatic-Ehren.f:100-116 (A = (rho + rho^T)/2, K = X o A - H' A),
 *                         :366-392 (calculate_rho)
 *
 * Matrices are column-major (Fortran order), leading dimension N.
 * Complex vectors are interleaved (re, im) double pairs.
 *
 * Pinning (tests/): SOLAP against ellipsoidal-coordinate quadrature and the
 * 1s-1s closed form; ROTAR against an independently built real-harmonic
 * rotation; the assembled S against invariants (symmetry, unit diagonal,
 * rotation/translation invariance of its spectrum); the Taylor propagator
 * against the exact generalised-eigen propagator (scipy.linalg.eigh); X_ij
 * against h_ij / S_ij; the Ehrenfest kernel against <K, S>_F = tr(hA) - tr(hA) = 0,
 * the symmetry of A and u^T A u = |u^T psi_e|^2 - |u^T psi_h|^2 for real u.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MXN 15
#define MXL 5
#define MXH (MXN + 3 * MXL)
#define LT MXL /* ltot of the rotation arrays */

/* Parallel regions only for large N: at small N thousands of tiny regions per
 * propagation cost more than the work (and stall behind other spinning pools). */
#define ORC_PAR_N 1024

static double re_[1001], reali_[1001], fact_[151], facti_[151];
static double root_[2 * MXL + 1], rooti_[2 * MXL + 1];
static int tables_ready = 0;

/* overlap_D.f:1194-1231 */
void orc_util_init(void) {
    if (tables_ready) return;
    fact_[0] = 1.0;
    facti_[0] = 1.0;
    for (int i = 1; i <= 150; ++i) {
        fact_[i] = fact_[i - 1] * i;
        facti_[i] = 1.0 / fact_[i];
    }
    re_[0] = 0.0;
    reali_[0] = 0.0;
    for (int i = 1; i <= 1000; ++i) {
        re_[i] = re_[i - 1] + 1.0;
        reali_[i] = 1.0 / re_[i];
    }
    root_[0] = 0.0;
    rooti_[0] = 1.0e200;
    for (int i = 1; i <= 2 * MXL; ++i) {
        root_[i] = sqrt(re_[i]);
        rooti_[i] = 1.0 / root_[i];
    }
    tables_ready = 1;
}

/* util.f:22-36 */
static double factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= (double)i;
    return f;
}

/* ------------------------------------------------------------------------ */
/* SOLAP: overlap_D.f:658-962.  Lined-up frame: A at the origin, B on +z.    */
/* sol[m], m = 0..min(la,lb), is <na la m | nb lb m> for unnormalised STOs.   */
/* ------------------------------------------------------------------------ */
#define H1(a, g) h1f1[(a) * (2 * (MXH + 1) + 1) + (g)]
#define SM(n, np) smat[(n) * (MXH + 1) + (np)]
#define SA(n, np) smataux[(n) * (MXH + 1) + (np)]
#define SMM(n, np) smatm[(n) * (MXH + 1) + (np)]

void orc_solap(int na, int la, double za, int nb, int lb, double zb, double R, double *sol) {
    static const double pi4 = 12.566370614359172953850573533118012;
    double h1f1[(MXH + 2) * (2 * (MXH + 1) + 1)];
    double smat[(MXH + 1) * (MXH + 1)], smataux[(MXH + 1) * (MXH + 1)], smatm[(MXH + 1) * (MXH + 1)];
    int nmin, nmax, npmin, npmax;
    double z, zp;
    orc_util_init();
    memset(h1f1, 0, sizeof h1f1);
    memset(smat, 0, sizeof smat);
    memset(smataux, 0, sizeof smataux);
    memset(smatm, 0, sizeof smatm);
    for (int m = 0; m <= MXL; ++m) sol[m] = 0.0;

    if (za >= zb) {
        nmin = na - la; nmax = na + la + lb + lb;
        npmin = nb - lb; npmax = nb + lb + la + la;
        z = za; zp = zb;
    } else {
        nmin = nb - lb; nmax = nb + lb + la + la;
        npmin = na - la; npmax = na + la + lb + lb;
        z = zb; zp = za;
    }
    const double expzR = exp(-z * R);

    /* confluent hypergeometric functions 1F1[a, g, arg] */
    const int ngmax = nmax + npmax + 2;
    const int namax = npmax + 1;
    const double arg = (z - zp) * R;
    const int iz = (int)z;
    int ia = (ngmax - iz) / 2;
    if (ia < 1) ia = 1;
    if (ia >= namax) ia = namax - 1;

    double auxa = arg * re_[ia] * reali_[ngmax];
    double f11a = 1.0;
    double auxb = arg * re_[ia + 1] * reali_[ngmax];
    double f11b = 1.0;
    for (int i = 1; i <= 1000; ++i) {
        f11a += auxa;
        f11b += auxb;
        auxa = auxa * re_[ia + i] * arg * reali_[ngmax + i] * reali_[i + 1];
        auxb = auxb * re_[ia + i + 1] * arg * reali_[ngmax + i] * reali_[i + 1];
        if (auxb < 1.0e-35 * f11b) break;
    }
    H1(ia, ngmax) = f11a;
    H1(ia + 1, ngmax) = f11b;

    double aux = arg - re_[ngmax];
    for (int i = ia + 1; i <= namax - 1; ++i)
        H1(i + 1, ngmax) = ((aux + re_[i + i]) * H1(i, ngmax) + re_[ngmax - i] * H1(i - 1, ngmax)) * reali_[i];
    for (int i = ia; i >= 2; --i)
        H1(i - 1, ngmax) = (re_[i] * H1(i + 1, ngmax) - (aux + re_[i + i]) * H1(i, ngmax)) * reali_[ngmax - i];

    for (int j = ngmax - 1; j >= 2; --j) {
        H1(1, j) = (H1(2, j + 1) + re_[j - 1] * H1(1, j + 1)) * reali_[j];
        aux = arg * reali_[j];
        int top = namax - 1 < j - 2 ? namax - 1 : j - 2;
        for (int i = 1; i <= top; ++i) H1(i + 1, j) = aux * H1(i + 1, j + 1) + H1(i, j);
    }

    /* basic integrals <N 0 0 | N' 0 0>, stored as smat(A index, B index) */
    {
        const double zzpR = (z + zp) * R;
        const double zzpi = 1.0 / (z + zp);
        double zzpin = pow(zzpi, nmin + npmin + 1);
        for (int n = nmin; n <= nmax; ++n) {
            double zzpinp = zzpin;
            for (int np = npmin; np <= npmax - n + nmin; ++np) {
                double auxk = fact_[n + np] * zzpinp;
                double sumk = 0.0;
                for (int k = 0; k <= n + np; ++k) {
                    double sumj = 0.0;
                    int j0 = k - n > 0 ? k - n : 0;
                    int j1 = k < np ? k : np;
                    for (int j = j0; j <= j1; ++j) sumj += H1(j + 1, k + 2) * facti_[n + j - k] * facti_[np - j];
                    sumk += sumj * auxk;
                    auxk = auxk * zzpR * reali_[n + np - k] * reali_[k + 2];
                }
                double v = pi4 * expzR * fact_[n] * fact_[np] * sumk;
                if (za >= zb) {
                    SM(n, np) = v; SA(n, np) = v; SMM(n, np) = v;
                } else {
                    SM(np, n) = v; SA(np, n) = v; SMM(np, n) = v;
                }
                zzpinp *= zzpi;
            }
            zzpin *= zzpi;
        }
    }

    nmin = na - la; nmax = na + la + lb + lb;
    npmin = nb - lb; npmax = nb + lb + la + la;
    const double Rinv = 1.0 / R, R2 = R * R, R4 = R2 * R2;
    double updltm0i = 0.5;
    const int mtop = la < lb ? la : lb;
    for (int m = 0; m <= mtop; ++m) {
        if (la > m) {
            for (int np = npmin; np <= npmax - 2; ++np) {
                double saux = SM(nmin, np);
                for (int n = nmin + 1; n <= nmax - 1 - np + npmin; ++n) {
                    double sbux = re_[m + m + 1] * 0.5 * Rinv * (SM(n + 1, np) + R2 * saux - SM(n - 1, np + 2));
                    saux = SM(n, np);
                    SA(n, np) = saux;
                    SM(n, np) = sbux;
                }
            }
            for (int ll = m + 1; ll <= la - 1; ++ll) {
                for (int np = npmin; np <= npmax - 2 * (ll + 1 - m); ++np) {
                    double saux = SM(nmin + ll - m, np);
                    for (int n = nmin + 1 + ll - m; n <= nmax - 1 - np + npmin - ll + m; ++n) {
                        double sbux = re_[ll + ll + 1] * 0.5 * reali_[ll - m + 1] * Rinv *
                                          (SM(n + 1, np) + R2 * saux - SM(n - 1, np + 2)) -
                                      re_[ll + m] * reali_[ll - m + 1] * SA(n, np);
                        saux = SM(n, np);
                        SA(n, np) = saux;
                        SM(n, np) = sbux;
                    }
                }
            }
        }
        if (lb > m) {
            for (int n = nmin + la - m; n <= nmax - la + m - 2; ++n) {
                double saux = SM(n, npmin);
                for (int np = npmin + 1; np <= npmax - 2 * (la - m) - n + nmin + la - m - 1; ++np) {
                    double sbux = -re_[m + m + 1] * 0.5 * Rinv * (SM(n, np + 1) + R2 * saux - SM(n + 2, np - 1));
                    saux = SM(n, np);
                    SA(n, np) = saux;
                    SM(n, np) = sbux;
                }
            }
            for (int ll = m + 1; ll <= lb - 1; ++ll) {
                for (int n = nmin + la - m; n <= nmax - la + m - 2 * (ll + 1 - m); ++n) {
                    double saux = SM(n, npmin + ll - m);
                    for (int np = npmin + 1 + ll - m; np <= npmax - 2 * (la - m) - n + nmin + la - m - ll + m - 1; ++np) {
                        double sbux = -re_[ll + ll + 1] * 0.5 * reali_[ll - m + 1] * Rinv *
                                          (SM(n, np + 1) + R2 * saux - SM(n + 2, np - 1)) -
                                      re_[ll + m] * reali_[ll - m + 1] * SA(n, np);
                        saux = SM(n, np);
                        SA(n, np) = saux;
                        SM(n, np) = sbux;
                    }
                }
            }
        }
        sol[m] = SM(na, nb);
        if (m == mtop) break;

        nmin += 1; nmax -= 3; npmin += 1; npmax -= 3;
        aux = updltm0i * 0.25 * re_[m + m + 1] * re_[m + m + 1] * Rinv * Rinv;
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
}
#undef H1
#undef SM
#undef SA
#undef SMM

/* overlap_D.f:466-478 */
double orc_anor_div(int la, int lb, int m) {
    double d = (la + m + 1.0) * (la - m) * (lb + m + 1.0) * (lb - m);
    return sqrt(d);
}

/* overlap_D.f:496-517 */
double orc_anor(double expa, double expb, int na, int nb, int la, int lb, double aux) {
    const double PI = 4.0 * atan(1.0);
    double anor = pow(expa * 2, na * 2 + 1);
    anor *= pow(expb * 2, nb * 2 + 1);
    anor *= (la * 2 + 1.0);
    anor *= (lb * 2 + 1.0);
    anor *= aux;
    return sqrt(anor) / (4 * PI);
}

/* ------------------------------------------------------------------------ */
/* ROTAR / DLMN: overlap_D.f:983-1191.  rl(m', m, l) real-harmonic rotation  */
/* matrices, -LT <= m', m <= LT, 0 <= l <= LT; column-major like Fortran.    */
/* ------------------------------------------------------------------------ */
#define RIX(mp, m, l) (((mp) + LT) + (2 * LT + 1) * (((m) + LT) + (2 * LT + 1) * (l)))

static void dlmn(int l, double sinal, double cosal, double cosbet, double tgbet2, double singa, double cosga,
                 double *dl, double *rl) {
    const double root2 = 1.414213562373095, one = 1.0, pt5 = 0.5;
    int iinf = 1 - l, isup = -iinf;
    dl[RIX(l, l, l)] = dl[RIX(isup, isup, l - 1)] * (one + cosbet) * pt5;
    dl[RIX(l, -l, l)] = dl[RIX(isup, -isup, l - 1)] * (one - cosbet) * pt5;
    for (int m = isup; m >= iinf; --m)
        dl[RIX(l, m, l)] = -tgbet2 * root_[l + m + 1] * rooti_[l - m] * dl[RIX(l, m + 1, l)];

    double al = l, al1 = al - one, tal1 = al + al1, ali = one / al1, cosaux = cosbet * al * al1;
    for (int mp = l - 1; mp >= 0; --mp) {
        double amp = mp;
        int laux = l + mp, lbux = l - mp;
        double aux = rooti_[laux] * rooti_[lbux] * ali;
        double cux = root_[laux - 1] * root_[lbux - 1] * al;
        for (int m = isup; m >= iinf; --m) {
            double am = m;
            int lauz = l + m, lbuz = l - m;
            double auz = rooti_[lauz] * rooti_[lbuz];
            double factor = aux * auz;
            double term = tal1 * (cosaux - am * amp) * dl[RIX(mp, m, l - 1)];
            if (lbuz != 1 && lbux != 1) {
                double cuz = root_[lauz - 1] * root_[lbuz - 1];
                term -= dl[RIX(mp, m, l - 2)] * cux * cuz;
            }
            dl[RIX(mp, m, l)] = factor * term;
        }
        iinf += 1;
        isup -= 1;
    }
    /* reflexion */
    double sign = one;
    iinf = -l;
    isup = l - 1;
    for (int m = l; m >= 1; --m) {
        for (int mp = iinf; mp <= isup; ++mp) {
            dl[RIX(mp, m, l)] = sign * dl[RIX(m, mp, l)];
            sign = -sign;
        }
        iinf += 1;
        isup -= 1;
    }
    /* inversion */
    iinf = -l;
    isup = iinf;
    for (int m = l - 1; m >= -l; --m) {
        sign = -one;
        for (int mp = isup; mp >= iinf; --mp) {
            dl[RIX(mp, m, l)] = sign * dl[RIX(-mp, -m, l)];
            sign = -sign;
        }
        isup += 1;
    }
    /* real harmonics */
    rl[RIX(0, 0, l)] = dl[RIX(0, 0, l)];
    double cosmal = cosal, sinmal = sinal;
    sign = -one;
    for (int mp = 1; mp <= l; ++mp) {
        double cosmga = cosga, sinmga = singa;
        double aux = root2 * dl[RIX(0, mp, l)];
        rl[RIX(mp, 0, l)] = aux * cosmal;
        rl[RIX(-mp, 0, l)] = aux * sinmal;
        for (int m = 1; m <= l; ++m) {
            aux = root2 * dl[RIX(m, 0, l)];
            rl[RIX(0, m, l)] = aux * cosmga;
            rl[RIX(0, -m, l)] = -aux * sinmga;
            double d1 = dl[RIX(-mp, -m, l)];
            double d2 = sign * dl[RIX(mp, -m, l)];
            double cosag = cosmal * cosmga - sinmal * sinmga;
            double cosagm = cosmal * cosmga + sinmal * sinmga;
            double sinag = sinmal * cosmga + cosmal * sinmga;
            double sinagm = sinmal * cosmga - cosmal * sinmga;
            rl[RIX(mp, m, l)] = d1 * cosag + d2 * cosagm;
            rl[RIX(mp, -m, l)] = -d1 * sinag + d2 * sinagm;
            rl[RIX(-mp, m, l)] = d1 * sinag + d2 * sinagm;
            rl[RIX(-mp, -m, l)] = d1 * cosag - d2 * cosagm;
            aux = cosmga * cosga - sinmga * singa;
            sinmga = sinmga * cosga + cosmga * singa;
            cosmga = aux;
        }
        sign = -sign;
        double aux2 = cosmal * cosal - sinmal * sinal;
        sinmal = sinmal * cosal + cosmal * sinal;
        cosmal = aux2;
    }
}

/* rl must hold (2LT+1)^2 (LT+1) doubles; layout RIX. */
void orc_rotar(int lmax, double cosal, double sinal, double cosbet, double sinbet, double *rl) {
    const double root2 = 1.414213562373095, zero = 0.0, one = 1.0, pt5 = 0.5;
    const double cosga = 1.0, singa = 0.0;
    double dl[(2 * LT + 1) * (2 * LT + 1) * (LT + 1)];
    orc_util_init();
    memset(dl, 0, sizeof dl);
    memset(rl, 0, sizeof(double) * (2 * LT + 1) * (2 * LT + 1) * (LT + 1));
    dl[RIX(0, 0, 0)] = one;
    rl[RIX(0, 0, 0)] = one;
    if (lmax == 0) return;
    dl[RIX(1, 1, 1)] = (one + cosbet) * pt5;
    dl[RIX(1, 0, 1)] = -sinbet / root2;
    dl[RIX(1, -1, 1)] = (one - cosbet) * pt5;
    dl[RIX(0, 1, 1)] = -dl[RIX(1, 0, 1)];
    dl[RIX(0, 0, 1)] = dl[RIX(1, 1, 1)] - dl[RIX(1, -1, 1)];
    dl[RIX(0, -1, 1)] = dl[RIX(1, 0, 1)];
    dl[RIX(-1, 1, 1)] = dl[RIX(1, -1, 1)];
    dl[RIX(-1, 0, 1)] = dl[RIX(0, 1, 1)];
    dl[RIX(-1, -1, 1)] = dl[RIX(1, 1, 1)];
    double cosag = cosal * cosga - sinal * singa;
    double cosamg = cosal * cosga + sinal * singa;
    double sinag = sinal * cosga + cosal * singa;
    double sinamg = sinal * cosga - cosal * singa;
    rl[RIX(0, 0, 1)] = dl[RIX(0, 0, 1)];
    rl[RIX(1, 0, 1)] = root2 * dl[RIX(0, 1, 1)] * cosal;
    rl[RIX(-1, 0, 1)] = root2 * dl[RIX(0, 1, 1)] * sinal;
    rl[RIX(0, 1, 1)] = root2 * dl[RIX(1, 0, 1)] * cosga;
    rl[RIX(0, -1, 1)] = -root2 * dl[RIX(1, 0, 1)] * singa;
    rl[RIX(1, 1, 1)] = dl[RIX(1, 1, 1)] * cosag - dl[RIX(1, -1, 1)] * cosamg;
    rl[RIX(1, -1, 1)] = -dl[RIX(1, 1, 1)] * sinag - dl[RIX(1, -1, 1)] * sinamg;
    rl[RIX(-1, 1, 1)] = dl[RIX(1, 1, 1)] * sinag - dl[RIX(1, -1, 1)] * sinamg;
    rl[RIX(-1, -1, 1)] = dl[RIX(1, 1, 1)] * cosag + dl[RIX(1, -1, 1)] * cosamg;

    double tgbet2;
    if (fabs(sinbet) < 1.0e-14) {
        tgbet2 = zero;
        sinbet = zero;
        cosbet = copysign(one, cosbet);
    } else if (fabs(sinbet) < 1.0e-10) {
        tgbet2 = zero;
        sinbet = zero;
        cosbet = copysign(one, cosbet);
    } else {
        tgbet2 = (one - cosbet) / sinbet;
    }
    for (int l = 2; l <= lmax; ++l) dlmn(l, sinal, cosal, cosbet, tgbet2, singa, cosga, dl, rl);
}

/* overlap_D.f:520-583: Euler angles of the A->B axis, gamma = 0. */
void orc_rotation_overlap(const double *xa, const double *xb, int lmax, double *Rab, double *rl) {
    const double tol = 1.0e-10;
    double xab = xb[0] - xa[0], yab = xb[1] - xa[1], zab = xb[2] - xa[2];
    double xy = sqrt(xab * xab + yab * yab);
    double sinal, cosal;
    if (xy > tol) {
        sinal = yab / xy;
        cosal = xab / xy;
    } else {
        sinal = 0.0;
        cosal = 1.0;
    }
    *Rab = sqrt(xab * xab + yab * yab + zab * zab);
    double sinbet = xy / *Rab, cosbet = zab / *Rab;
    orc_rotar(lmax, cosal, sinal, cosbet, sinbet, rl);
}

/* ------------------------------------------------------------------------ */
/* Build_Overlap_Matrix: overlap_D.f:144-311 (no PBC, all atoms QM).         */
/*   coords: n_atoms x 3 row-major (x, y, z per atom), Angstrom             */
/*   first_orb[ia] .. first_orb[ia+1]-1: orbitals of atom ia                */
/*   atom_lmax[ia]: AngMax of atom ia's element                             */
/*   orbital k: n, l, m, nzeta, zeta[2k..2k+1] (1/Angstrom), coef[2k..2k+1]  */
/*   drop_below > 0 zeroes off-diagonal |S| below it, as the reference's    */
/*   unmoved-atom recycle cache does (overlap_D.f:193-225, SPRSIN :586-655). */
/* S: N x N column-major, symmetric, unit diagonal.                          */
/* ------------------------------------------------------------------------ */
int orc_overlap(int n_atoms, const double *coords, const int *first_orb, const int *atom_lmax, int n_orb,
                const int *on, const int *ol, const int *om, const int *onz, const double *ozeta,
                const double *ocoef, double cutoff, double drop_below, double *S) {
    orc_util_init();
    const long N = n_orb;
    memset(S, 0, sizeof(double) * N * N);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < n_atoms; ++ib) {
        double rl[(2 * LT + 1) * (2 * LT + 1) * (LT + 1)];
        double solvec[MXL + 1], solnorm[MXL + 1], sol_partial[MXL + 1], sux[11];
        for (int ia = ib + 1; ia < n_atoms; ++ia) {
            double Rab;
            int lmax = atom_lmax[ia] > atom_lmax[ib] ? atom_lmax[ia] : atom_lmax[ib];
            orc_rotation_overlap(&coords[3 * ia], &coords[3 * ib], lmax, &Rab, rl);
            if (Rab > cutoff) continue;
            for (int b = first_orb[ib]; b < first_orb[ib + 1]; ++b) {
                for (int a = first_orb[ia]; a < first_orb[ia + 1]; ++a) {
                    int nb = on[b], lb = ol[b], mb = om[b];
                    int na = on[a], la = ol[a], ma = om[a];
                    double aux = 1.0 / (factorial(na + na) * factorial(nb + nb));
                    int msup = la < lb ? la : lb;
                    for (int k = 0; k <= msup; ++k) solnorm[k] = 0.0;
                    for (int j = 0; j < onz[b]; ++j) {
                        for (int i = 0; i < onz[a]; ++i) {
                            double expb = ozeta[2 * b + j], expa = ozeta[2 * a + i];
                            orc_solap(na, la, expa, nb, lb, expb, Rab, solvec);
                            double anor = orc_anor(expa, expb, na, nb, la, lb, aux);
                            for (int m = 0; m <= msup - 1; ++m) {
                                sol_partial[m] = solvec[m] * anor;
                                anor = anor / orc_anor_div(la, lb, m);
                                if (m == 0) anor = anor + anor;
                            }
                            sol_partial[msup] = solvec[msup] * anor;
                            for (int k = 0; k <= msup; ++k)
                                solnorm[k] += ocoef[2 * a + i] * ocoef[2 * b + j] * sol_partial[k];
                        }
                    }
                    sux[0] = solnorm[0] * rl[RIX(ma, 0, la)] * rl[RIX(mb, 0, lb)];
                    for (int k = 1; k <= msup; ++k)
                        sux[k] = solnorm[k] * (rl[RIX(ma, -k, la)] * rl[RIX(mb, -k, lb)] +
                                               rl[RIX(ma, k, la)] * rl[RIX(mb, k, lb)]);
                    double s = 0.0;
                    for (int k = 0; k <= msup; ++k) s += sux[k];
                    S[(long)a + N * b] += s;
                }
            }
        }
    }
    for (long i = 0; i < N; ++i) S[i + N * i] = 1.0;
    if (drop_below > 0.0)
        for (long j = 0; j < N; ++j)
            for (long i = j + 1; i < N; ++i)
                if (fabs(S[i + N * j]) < drop_below) S[i + N * j] = 0.0;
    /* mirror the lower triangle into the upper one */
    for (long j = 0; j < N; ++j)
        for (long i = j + 1; i < N; ++i) S[j + N * i] = S[i + N * j];
    return 0;
}

/* X_ij * S_ij: hamiltonians.f:33-63 and ElHl_Chebyshev.f:296-323 */
/* X_ij of hamiltonians.f:33-63 */
static double huckel_x_ij(long i, long j, const double *ip, const double *kwh, const double *vshift) {
    if (i == j) return ip[i] + vshift[i];
    double c1 = ip[i] - ip[j];
    double c2 = ip[i] + ip[j];
    double c3 = (c1 / c2) * (c1 / c2);
    double c4 = (vshift[i] + vshift[j]) * 0.5;
    double k = (kwh[i] + kwh[j]) * 0.5;
    double keff = k + c3 + c3 * c3 * (1.0 - k);
    return keff * c2 * 0.5 + c4;
}

void orc_huckel(int n_orb, const double *ip, const double *kwh, const double *vshift, const double *S, double *h) {
    const long N = n_orb;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j) {
        for (long i = 0; i <= j; ++i) {
            const double x = huckel_x_ij(i, j, ip, kwh, vshift);
            h[i + N * j] = x * S[i + N * j];
            h[j + N * i] = h[i + N * j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Symmetric inverse.  The reference factors with dsytrf/dsytri (Bunch-      */
/* Kaufman) and symmetrises (GPU_Interface.cpp:936-949, Matrix_math.f:126-  */
/* 198).  S is SPD, so a Cholesky-based inverse computes the same matrix;    */
/* only the rounding differs.  Returns nonzero when S is not positive.       */
/* ------------------------------------------------------------------------ */
int orc_cholesky(int n, const double *A, double *L) {
    const long N = n;
    memcpy(L, A, sizeof(double) * N * N);
    for (long j = 0; j < N; ++j) {
        double d = L[j + N * j];
        for (long k = 0; k < j; ++k) d -= L[j + N * k] * L[j + N * k];
        if (!(d > 0.0)) return (int)(j + 1);
        d = sqrt(d);
        L[j + N * j] = d;
#pragma omp parallel for schedule(static) if (N >= ORC_PAR_N)
        for (long i = j + 1; i < N; ++i) {
            double s = L[i + N * j];
            for (long k = 0; k < j; ++k) s -= L[i + N * k] * L[j + N * k];
            L[i + N * j] = s / d;
        }
    }
    for (long j = 0; j < N; ++j)
        for (long i = 0; i < j; ++i) L[i + N * j] = 0.0;
    return 0;
}

int orc_syinvert(int n, const double *S, double *Sinv) {
    const long N = n;
    double *L = (double *)malloc(sizeof(double) * N * N);
    double *Li = (double *)calloc((size_t)(N * N), sizeof(double));
    int info = orc_cholesky(n, S, L);
    if (info) {
        free(L);
        free(Li);
        return info;
    }
    /* Li = L^-1 (lower), column by column */
#pragma omp parallel for schedule(dynamic, 8)
    for (long j = 0; j < N; ++j) {
        Li[j + N * j] = 1.0 / L[j + N * j];
        for (long i = j + 1; i < N; ++i) {
            double s = 0.0;
            for (long k = j; k < i; ++k) s -= L[i + N * k] * Li[k + N * j];
            Li[i + N * j] = s / L[i + N * i];
        }
    }
    /* S^-1 = Li^T Li, then symmetrise */
#pragma omp parallel for schedule(dynamic, 8)
    for (long j = 0; j < N; ++j)
        for (long i = 0; i <= j; ++i) {
            double s = 0.0;
            for (long k = j; k < N; ++k) s += Li[k + N * i] * Li[k + N * j];
            Sinv[i + N * j] = s;
            Sinv[j + N * i] = s;
        }
    free(L);
    free(Li);
    return 0;
}

/* C = A * B, all N x N column-major (xPU_dsymm with a full symmetric A). */
void orc_dgemm_nn(int n, const double *A, const double *B, double *C) {
    const long N = n;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j) {
        double *c = &C[N * j];
        for (long i = 0; i < N; ++i) c[i] = 0.0;
        for (long k = 0; k < N; ++k) {
            double b = B[k + N * j];
            const double *a = &A[N * k];
            for (long i = 0; i < N; ++i) c[i] += a[i] * b;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Taylor propagation: Taylor.f:35-303.                                      */
/* ------------------------------------------------------------------------ */
typedef double complex zc;

static const double H_BAR = 6.58264e-4; /* constants_m.f:23, eV*ps */

/* The reference loops without bound when the norm test cannot pass (a norm
 * tolerance below the accumulated rounding drift); the restatement gives up
 * after this many tau shrinks/rescales in one call and reports failure. */
#define ORC_GUARD 400

/* res = alpha * H' * ket  (Matrix_math.f:291-301, dzgemv 'N') */
static void op_x_ket(long N, zc *res, const double *H, const zc *ket, zc alpha) {
#pragma omp parallel if (N >= ORC_PAR_N)
    {
#pragma omp for schedule(static)
        for (long i = 0; i < N; ++i) res[i] = 0.0;
        /* row blocks per thread, column sweep: streams H once */
#pragma omp for schedule(static)
        for (long ib = 0; ib < N; ib += 64) {
            long ie = ib + 64 < N ? ib + 64 : N;
            zc acc[64];
            for (long i = ib; i < ie; ++i) acc[i - ib] = 0.0;
            for (long j = 0; j < N; ++j) {
                const double *h = &H[N * j];
                zc x = ket[j];
                double xr = creal(x), xi = cimag(x);
                for (long i = ib; i < ie; ++i) acc[i - ib] += h[i] * xr + I * (h[i] * xi);
            }
            for (long i = ib; i < ie; ++i) res[i] = alpha * acc[i - ib];
        }
    }
}

/* res = alpha * H'^T * bra  (Matrix_math.f:238-248, dzgemv 'T') */
static void bra_x_op(long N, zc *res, const zc *bra, const double *H, zc alpha) {
#pragma omp parallel for schedule(static) if (N >= ORC_PAR_N)
    for (long j = 0; j < N; ++j) {
        const double *h = &H[N * j];
        double sr = 0.0, si = 0.0;
        for (long i = 0; i < N; ++i) {
            sr += h[i] * creal(bra[i]);
            si += h[i] * cimag(bra[i]);
        }
        res[j] = alpha * (sr + I * si);
    }
}

static double dotc_abs(long N, const zc *a, const zc *b) {
    double sr = 0.0, si = 0.0;
#pragma omp parallel for reduction(+ : sr, si) schedule(static) if (N >= ORC_PAR_N)
    for (long i = 0; i < N; ++i) {
        zc p = conj(a[i]) * b[i];
        sr += creal(p);
        si += cimag(p);
    }
    return cabs(sr + I * si);
}

static int is_converged(long N, const zc *a, const zc *b, double tol) {
    for (long i = 0; i < N; ++i)
        if (cabs(a[i] - b[i]) > tol) return 0;
    return 1;
}

static void coefficient(double tau, int kmax, zc *c) {
    c[0] = 1.0;
    for (int k = 2; k <= kmax; ++k) c[k - 1] = -I * c[k - 2] * (tau / (k - 1));
}

typedef struct {
    long n_matvec;   /* GEMVs performed (bra and ket counted separately) */
    long n_substeps; /* accepted sub-steps of the while loop + convergence steps */
    long n_rescale;  /* norm-failure rescales */
    long n_shrink;   /* convergence-failure shrinks */
} orc_stats;

static int convergence(long N, zc *psi_bra, zc *psi_ket, zc *C, int *k_ref, double tau, const double *Hp,
                       double norm_ref, int order, double err_tol, double norm_tol, zc *bra, zc *ket, zc *tb,
                       zc *tk, orc_stats *st) {
    zc *tb_old = tb, *tb_new = tb + N, *tk_old = tk, *tk_new = tk + N;
    coefficient(tau, order, C);
    int kmax = order;
    for (int k = 2; k <= order; ++k)
        if (cabs(C[k - 1]) < 1.0e-16) {
            kmax = k;
            break;
        }
    *k_ref = kmax;
    memcpy(bra, psi_bra, sizeof(zc) * N);
    memcpy(ket, psi_ket, sizeof(zc) * N);
    memcpy(tb_old, bra, sizeof(zc) * N);
    memcpy(tk_old, ket, sizeof(zc) * N);
    for (int k = 2; k <= kmax; ++k) {
        zc r = C[k - 1] / C[k - 2];
        bra_x_op(N, bra + N * (k - 1), bra + N * (k - 2), Hp, r);
        op_x_ket(N, ket + N * (k - 1), Hp, ket + N * (k - 2), r);
        st->n_matvec += 2;
        for (long i = 0; i < N; ++i) {
            tb_new[i] = tb_old[i] + bra[N * (k - 1) + i];
            tk_new[i] = tk_old[i] + ket[N * (k - 1) + i];
        }
        if (is_converged(N, tb_new, tb_old, err_tol) && is_converged(N, tk_new, tk_old, err_tol)) {
            double norm_tmp = dotc_abs(N, tb_new, tk_new);
            if (fabs(norm_tmp - norm_ref) < norm_tol) {
                memcpy(psi_bra, tb_new, sizeof(zc) * N);
                memcpy(psi_ket, tk_new, sizeof(zc) * N);
                st->n_substeps += 1;
                return 1;
            }
        }
        memcpy(tb_old, tb_new, sizeof(zc) * N);
        memcpy(tk_old, tk_new, sizeof(zc) * N);
    }
    return 0;
}

/*
 * Propagation (Taylor.f:35-127).  bra, ket: N complex (interleaved) in/out.
 * tau in/out, save_tau out.  err_tol / norm_tol: the reference uses 1e-8
 * (Taylor.f:21-22); order: 25 (Taylor.f:20).  stats: 4 longs out (may be NULL).
 */
int orc_propagation(int n, const double *Hp, double *bra_io, double *ket_io, double t_init, double t_max,
                    double *tau_io, double *save_tau, double err_tol, double norm_tol, int order, long *stats) {
    const long N = n;
    zc *psi_bra = (zc *)bra_io, *psi_ket = (zc *)ket_io;
    zc *bra = (zc *)malloc(sizeof(zc) * N * order);
    zc *ket = (zc *)malloc(sizeof(zc) * N * order);
    zc *tb = (zc *)malloc(sizeof(zc) * N * 2);
    zc *tk = (zc *)malloc(sizeof(zc) * N * 2);
    zc *C = (zc *)malloc(sizeof(zc) * order);
    orc_stats st = {0, 0, 0, 0};
    double tau = *tau_io;
    int k_ref = order;

    double norm_ref = dotc_abs(N, psi_bra, psi_ket);
    int guard = 0;
    while (!convergence(N, psi_bra, psi_ket, C, &k_ref, tau, Hp, norm_ref, order, err_tol, norm_tol, bra, ket,
                        tb, tk, &st)) {
        tau *= 0.9;
        st.n_shrink++;
        if (++guard > ORC_GUARD) break;
    }
    *save_tau = tau;
    double t = t_init + tau * H_BAR;
    if (t_max - t < tau * H_BAR) {
        tau = (t_max - t) / H_BAR;
        coefficient(tau, order, C);
    }
    while (t < t_max) {
        memcpy(bra, psi_bra, sizeof(zc) * N);
        memcpy(ket, psi_ket, sizeof(zc) * N);
        zc *tmp_bra = tb, *tmp_ket = tk;
        memcpy(tmp_bra, bra, sizeof(zc) * N);
        memcpy(tmp_ket, ket, sizeof(zc) * N);
        for (int k = 2; k <= k_ref; ++k) {
            zc r = C[k - 1] / C[k - 2];
            bra_x_op(N, bra + N * (k - 1), bra + N * (k - 2), Hp, r);
            op_x_ket(N, ket + N * (k - 1), Hp, ket + N * (k - 2), r);
            st.n_matvec += 2;
            for (long i = 0; i < N; ++i) {
                tmp_bra[i] += bra[N * (k - 1) + i];
                tmp_ket[i] += ket[N * (k - 1) + i];
            }
        }
        double norm_test = dotc_abs(N, tmp_bra, tmp_ket);
        if (fabs(norm_test - norm_ref) < norm_tol) {
            memcpy(psi_bra, tmp_bra, sizeof(zc) * N);
            memcpy(psi_ket, tmp_ket, sizeof(zc) * N);
            st.n_substeps++;
        } else {
            int ok = 0;
            while (!ok) {
                tau *= 0.975;
                st.n_rescale++;
                ok = convergence(N, psi_bra, psi_ket, C, &k_ref, tau, Hp, norm_ref, order, err_tol, norm_tol, bra,
                                 ket, tb, tk, &st);
                if (++guard > ORC_GUARD) break;
            }
        }
        t += tau * H_BAR;
        if (t_max - t < tau * H_BAR) {
            tau = (t_max - t) / H_BAR;
            coefficient(tau, order, C);
        }
    }
    *tau_io = tau;
    if (stats) {
        stats[0] = st.n_matvec;
        stats[1] = st.n_substeps;
        stats[2] = st.n_rescale;
        stats[3] = st.n_shrink;
    }
    free(bra);
    free(ket);
    free(tb);
    free(tk);
    free(C);
    return guard > ORC_GUARD ? -1 : 0;
}

/* OpenMP thread count of the following calls (the CPU baseline runs all cores and one). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

/* Bare GEMV kernels of the Taylor path, exported for the CPU baseline. */
void orc_op_x_ket(int n, const double *H, const double *ket, double *res) {
    op_x_ket(n, (zc *)res, H, (const zc *)ket, 1.0);
}
void orc_bra_x_op(int n, const double *H, const double *bra, double *res) {
    bra_x_op(n, (zc *)res, (const zc *)bra, H, 1.0);
}

/*
 * Populations_mtx / pop_Slater (data_output.f:87-147,242-263) for one
 * particle: pop[f] = Re sum_{i in fragment f} dual_bra_i * dual_ket_i for
 * f < n_frag, pop[n_frag] = total.  frag[i] in [0, n_frag).
 */
void orc_populations(int n, int n_frag, const int *frag, const double *dual_bra, const double *dual_ket,
                     double *pop) {
    const zc *a = (const zc *)dual_bra, *b = (const zc *)dual_ket;
    for (int f = 0; f <= n_frag; ++f) pop[f] = 0.0;
    zc tot = 0.0;
    zc *acc = (zc *)calloc((size_t)n_frag, sizeof(zc));
    for (long i = 0; i < n; ++i) {
        zc p = a[i] * b[i];
        acc[frag[i]] += p;
        tot += p;
    }
    for (int f = 0; f < n_frag; ++f) pop[f] = creal(acc[f]);
    pop[n_frag] = creal(tot);
    free(acc);
}

/* QuasiParticleEnergies (ElHl_Chebyshev.f:329-371): sum_ij AO_bra_i H_ij AO_ket_j */
void orc_qp_energy(int n, const double *ao_bra, const double *H, const double *ao_ket, double *erg) {
    const long N = n;
    const zc *a = (const zc *)ao_bra, *b = (const zc *)ao_ket;
    double er = 0.0, ei = 0.0;
#pragma omp parallel for reduction(+ : er, ei) schedule(static)
    for (long j = 0; j < N; ++j) {
        zc s = 0.0;
        for (long i = 0; i < N; ++i) s += a[i] * H[i + N * j];
        zc p = s * b[j];
        er += creal(p);
        ei += cimag(p);
    }
    erg[0] = er;
    erg[1] = ei;
}

/* Huckel_stuff (diabatic-Ehren.f:250-276): the full symmetric X_ij matrix. */
void orc_huckel_x(int n_orb, const double *ip, const double *kwh, const double *vshift, double *X) {
    const long N = n_orb;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j)
        for (long i = 0; i <= j; ++i) {
            X[i + N * j] = huckel_x_ij(i, j, ip, kwh, vshift);
            X[j + N * i] = X[i + N * j];
        }
}

/*
 * Diabatic-Ehrenfest kernel matrix, CPU branch of Diabatic_Ehrenfest
 * (diabatic-Ehren.f:100-116):
 *   rho_ij = Re(ket_j,1 bra_i,1) - Re(ket_j,2 bra_i,2)   calculate_rho (:366-392)
 *   A      = (rho + rho^T) / two                         (:104)
 *   K      = X * A - matmul(H', A)                       (:112-116)
 * bra, ket: AO_bra(N, n_part), AO_ket(N, n_part), complex column-major, bra as
 * the caller holds it (conjugated, ElHl_Chebyshev.f:277-279); n_part = 1 drops
 * the hole term.  A (may be NULL) receives A.
 */
void orc_ehrenfest_kernel(int n, int n_part, const double *Hp, const double *X, const double *bra,
                          const double *ket, double *A_out, double *K) {
    const long N = n;
    double *rho = malloc(sizeof(double) * N * N);
    double *A = malloc(sizeof(double) * N * N);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j)
        for (long i = 0; i < N; ++i) {
            double r = 0.0;
            for (int p = 0; p < n_part; ++p) {
                const double *kj = &ket[2 * (j + N * p)], *bi = &bra[2 * (i + N * p)];
                const double v = kj[0] * bi[0] - kj[1] * bi[1];
                r = p == 0 ? v : r - v;
            }
            rho[i + N * j] = r;
        }
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j)
        for (long i = 0; i < N; ++i) A[i + N * j] = (rho[i + N * j] + rho[j + N * i]) / 2.0;
    orc_dgemm_nn(n, Hp, A, K);  /* B = H' A */
#pragma omp parallel for schedule(static)
    for (long j = 0; j < N; ++j)
        for (long i = 0; i < N; ++i) K[i + N * j] = X[i + N * j] * A[i + N * j] - K[i + N * j];
    if (A_out) memcpy(A_out, A, sizeof(double) * N * N);
    free(rho);
    free(A);
}
