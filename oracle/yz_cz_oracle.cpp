/*
 * yz_cz_oracle.cpp -- CPU FP64 restatement of Yuzen's ConcZ step: a 2-D
 * film whose concentration is resolved through the thickness.
 *
 * TEST INFRASTRUCTURE ONLY: compiled into liboracle.so with yz_oracle.cpp
 * (it is included at that file's end and uses its CRS, BiCGSTAB and
 * multigrid); the product never includes, links or calls it.
 *
 * It restates, in plain loops over the reference's own row layout (two
 * ghost layers on every side of the 2-D film grid and of the 3-D
 * concentration grid; row = Nx iy + ix on the film, Nx Ny kz + Nx iy + ix
 * on the concentration), the functions below. Paths are relative to
 * yuzen/src/model/ConcZ/CZ2D/ unless given in full.
 *
 *   face mobility, 1/mu means,  CZ_HeightEquation_2D.cpp:624-748
 *   surface concentration,
 *   Marangoni flux, wall faces
 *   film assembly (A, B, the   CZ_HeightEquation_2D.cpp:750-1081
 *   evaporation and its        (yuzen/src/model/Evaporation.cpp:62-220,
 *   solvent limit)              Evaporation.h:112-119)
 *   ghost rows of the film     CZ_HeightEquation_2D.cpp:1083-1142
 *   face flux parts, the       CZ_HeightEquation_2D.cpp:1161-1301
 *   conservative update
 *   the velocity through the   CZ_HeightEquation_2D.cpp:1381-1749 (the
 *   film (Simpson's rule)      in-plane components; the vertical one is
 *                              an output the concentration does not read)
 *   predictor + midpoint stage CZ_HeightEquation_2D.cpp:1949-2020,
 *                              yuzen/src/numerics/SemiImplicit.cpp:103-190
 *   coupling passes (Aitken),  CZ_HeightEquation_2D.cpp:207-386
 *   the step's completion
 *   RedistributeExcess,        CZ_HeightEquation_2D.cpp:388-570
 *   RedistributeDeficit
 *   volume fluxes (layer       CZ_ConcEquation_2D.cpp:144-263
 *   means, the geometric
 *   conservation law)
 *   the concentration rows     CZ_ConcEquation_2D.cpp:290-570
 *   the conormal ghost rows    CZ_ConcEquation_2D.cpp:588-732
 *   the concentration stage,   CZ_ConcEquation_2D.cpp:755-1005
 *   its solute correction,
 *   ExplicitPart, the ghosts
 *   properties                 yuzen/src/model/FluidProperty.h:110-183
 *
 * Both linear systems go through the oracle's backend restatement (row
 * scaling, BiCGSTAB with the smoothed-aggregation multigrid) to the
 * reference's default tolerance 1e-13; the reference's own backend is
 * AMGCL's (an un-vendored library), so the solution is pinned to that
 * tolerance and the iteration counts are not.
 *
 * The limited concentration schemes (Koren, van Leer) freeze their
 * limiter at the field the stage is solved at (AdvWeights :575-593) and
 * settle it to the midpoint of the step (SettleLimiter :945-973).
 *
 * Not restated (the step refuses them): the 1-D film (Ny = 1) and its
 * sideways outflow, step control, the manufactured solution, the
 * negative-power evaporation profile.
 */

namespace yzo {

struct CzStepper {
    yz_cz_case cs;
    int Nx = 0, Ny = 0, Nz = 0, n2 = 0, n3 = 0;
    double dx = 1, dy = 1, dz = 1, xlo = 0, ylo = 0;
    double dt = 0;
    const double theta = 0.5;
    double c1w[4], c2w[4];
    bool clim = false;           /* a limited scheme (Koren, van Leer) */
    Limiter lim;
    long long settle_passes = 0; /* SettleLimiter's solves over the run */
    bool wall[4];

    vector<double> S;                       /* substrate */
    vector<double> h_old, h_new;            /* film equation: _phi_old, _phi_new */
    vector<double> fH, fC;                  /* the field */
    vector<double> tH, tC;                  /* _field_tmp */
    vector<double> c_old, c_new;            /* concentration equation */
    vector<double> ccPrev; double dtPrev = 1.0; bool ccPrevSet = false;

    /* film system: 13 values per row, slot order as stored (the diagonal
     * first, then the reference's insertion order) */
    vector<double> Ah, Bh;                  /* A of the stage, B */
    vector<double> qCapX, qExpX, qCapY, qExpY, qX, qY, colSrc, dtdh;

    /* concentration */
    vector<double> Ux, Uy, Qx, Qy, W, react, Ac, Bc, cRhs;

    double st_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int last_iters = 0;
    double last_resid = 0;

    /* ---- indices ---- */
    int I2(int i, int j) const { return j * Nx + i; }
    int I3(int i, int j, int k) const { return k * Nx * Ny + j * Nx + i; }
    bool realx(int i) const { return i >= 2 && i <= Nx - 3; }
    bool realy(int j) const { return j >= 2 && j <= Ny - 3; }
    bool realz(int k) const { return k >= 2 && k <= Nz - 3; }

    /* ---- properties (FluidProperty.h:126-131, 178-183) ---- */
    double visc(double c) const { return 1.0 * std::exp(cs.visc_exp * (c - cs.visc_c0)); }
    double diffu(double c) const { return 1.0 * std::exp(cs.diff_exp * (c - cs.diff_c0)); }
    double sigma(double c) const { return 1.0 + cs.Ma * (c - cs.sig_c0); }

    /* ---- evaporation (Evaporation.cpp:62-103, 148-220) ---- */
    double evap_eval(double hh, double cc) const {
        double f_conc;
        if (cs.evap_func == YZ_EVAP_POWERN) {
            double v = 1.0 - cc - cs.evap_concr;
            v = std::max(v, 0.0);
            f_conc = std::pow(v, cs.evap_power);
        } else {
            const double c = (cc < 0.0) ? 0.0 : ((cc > 1.0) ? 1.0 : cc);
            const double chi = (c / cs.molar_resin) / (c / cs.molar_resin + (1.0 - c) / cs.molar_solvent);
            f_conc = 1.0 - chi;
        }
        double f_mod = 1.0;
        if (cs.precursor != 0.0) {
            const double v = (hh - cs.precursor) / ((20.0 - 1.0) * cs.precursor);
            f_mod = std::max(0.0, std::min(v, 1.0));
        }
        return cs.evap_coef * f_conc * 1.0 * f_mod;
    }
    /* EvapRate of the field at column (ix, jy): the surface value is the
     * mean of the top real cell and the ghost above it */
    double evap_rate(const vector<double> &H, const vector<double> &C, int ix, int jy) const {
        if (!cs.evaporation) return 0.0;
        const double cc = 0.5 * (C[I3(ix, jy, Nz - 3)] + C[I3(ix, jy, Nz - 2)]);
        return evap_eval(H[I2(ix, jy)], cc);
    }
    static double solvent_limit(double evp, double h, double c, double dt) {
        if (!(evp > 0.0)) return evp;
        const double cc = (c < 0.0) ? 0.0 : ((c > 1.0) ? 1.0 : c);
        const double solvent = (h > 0.0) ? h * (1.0 - cc) : 0.0;
        if (evp * dt > solvent) evp = solvent / dt;
        return evp;
    }
    double column_mean(const vector<double> &C, int ix, int jy) const {
        double s = 0.0;
        for (int kz = 2; kz < Nz - 2; kz++) s += C[I3(ix, jy, kz)];
        return s / (double)(Nz - 4);
    }

    /* ---- faces (CZ_HeightEquation_2D.cpp:624-748); dir 0: x face ix
     * between ix-1 and ix, 1: y face jy between jy-1 and jy ---- */
    bool wall_face(int ix, int jy, int dir) const {
        if (dir == 0) return (ix == 2 && wall[0]) || (ix == Nx - 2 && wall[1]);
        return (jy == 2 && wall[2]) || (jy == Ny - 2 && wall[3]);
    }
    double face_inv_visc(int ix, int jy, int dir, int power) const {
        const int ax = dir == 0, ay = dir == 1;
        double num = 0, den = 0;
        for (int kz = 2; kz < Nz - 2; kz++) {
            const double c = 0.5 * (fC[I3(ix - ax, jy - ay, kz)] + fC[I3(ix, jy, kz)]);
            const double r = 1.0 - dz * (-1.5 + (double)kz);
            const double w = (power == 2) ? r * r : r;
            num += w / visc(c);
            den += w;
        }
        return num / den;
    }
    double face_surf_conc(int ix, int jy, int dir) const {
        const int ax = dir == 0, ay = dir == 1;
        return 0.25 * (fC[I3(ix - ax, jy - ay, Nz - 2)] + fC[I3(ix - ax, jy - ay, Nz - 3)]
                       + fC[I3(ix, jy, Nz - 2)] + fC[I3(ix, jy, Nz - 3)]);
    }
    double face_mobility(int ix, int jy, int dir) const {
        if (wall_face(ix, jy, dir)) return 0.0;
        const int ax = dir == 0, ay = dir == 1;
        const double h1 = fH[I2(ix - ax, jy - ay)], h2 = fH[I2(ix, jy)];
        const double denom = h1 + h2;
        if (denom <= 0.0) return 0.0;
        const double mob = 2.0 * (h1 * h1) * (h2 * h2) / denom;
        return mob / 3.0 * face_inv_visc(ix, jy, dir, 2);
    }
    double marangoni_flux(int ix, int jy, int dir) const {
        if (wall_face(ix, jy, dir)) return 0.0;
        const int ax = dir == 0, ay = dir == 1;
        const double h1 = fH[I2(ix - ax, jy - ay)], h2 = fH[I2(ix, jy)];
        if (h1 + h2 <= 0.0) return 0.0;
        const double ds = dir == 0 ? dx : dy;
        const double cs1 = 0.5 * (fC[I3(ix - ax, jy - ay, Nz - 2)] + fC[I3(ix - ax, jy - ay, Nz - 3)]);
        const double cs2 = 0.5 * (fC[I3(ix, jy, Nz - 2)] + fC[I3(ix, jy, Nz - 3)]);
        const double dcds = (cs2 - cs1) / ds;
        const double h2eff = h1 * h2;
        return 0.5 * h2eff * cs.As * cs.Ma * dcds * face_inv_visc(ix, jy, dir, 1);
    }
    double xface(int ix) const { return xlo + ix * dx; }
    double yface(int jy) const { return ylo + jy * dy; }

    /* ---- the film's ghost rows (CZ_HeightEquation_2D.cpp:1083-1142) ---- */
    int fold(int i, int np, bool per, bool lo) const {
        const int m = np - 4;
        if (lo) return per ? i + m : 3 - i;
        return per ? i - m : 2 * (np - 3) + 1 - i;
    }
    /* the row a ghost row of the film copies (-1: not a ghost row) */
    int h_src(int ix, int jy) const {
        if (ix < 2 || ix > Nx - 3) return I2(fold(ix, Nx, !wall[0], ix < 2), jy);
        if (jy < 2 || jy > Ny - 3) return I2(ix, fold(jy, Ny, !wall[2], jy < 2));
        return -1;
    }
    /* CopyGhosts: the rows swept twice in row order */
    void copy_ghosts(vector<double> &h) const {
        for (int sweep = 0; sweep < 2; sweep++)
            for (int r = 0; r < n2; r++) {
                const int s = h_src(r % Nx, r / Nx);
                if (s < 0) continue;
                double sum = 0.0;
                sum -= (-1.0) * h[s];
                h[r] = sum / 1.0;
            }
    }

    /* slot order of a film row as stored: diagonal, then (-2,0) (-1,0)
     * (1,0) (2,0) (0,-2) (0,-1) (0,1) (0,2) (-1,-1) (-1,1) (1,-1) (1,1) */
    static constexpr int hox[13] = {0, -2, -1, 1, 2, 0, 0, 0, 0, -1, -1, 1, 1};
    static constexpr int hoy[13] = {0, 0, 0, 0, 0, -2, -1, 1, 2, -1, 1, -1, 1};

    /* ---- SetMatrixA of the film (:750-1081) + StoreFaceFluxParts ---- */
    void h_set_matrix() {
        std::fill(Ah.begin(), Ah.end(), 0.0);
        std::fill(Bh.begin(), Bh.end(), 0.0);
        const double rdx = 1.0 / dx, rdx2 = rdx * rdx, rdx4 = rdx2 * rdx2;
        const double rdy = 1.0 / dy, rdy2 = rdy * rdy, rdy4 = rdy2 * rdy2;
        const double rdx2dy2 = rdx2 * rdy2;
        auto SS = [&](int i, int j) { return S[I2(i, j)]; };
        auto HH = [&](int i, int j) { return fH[I2(i, j)]; };
        for (int jy = 2; jy < Ny - 2; jy++)
            for (int ix = 2; ix < Nx - 2; ix++) {
                double *a = &Ah[(size_t)I2(ix, jy) * 13];
                const double DL = face_mobility(ix, jy, 0), DR = face_mobility(ix + 1, jy, 0);
                const double sigL = sigma(face_surf_conc(ix, jy, 0)), sigR = sigma(face_surf_conc(ix + 1, jy, 0));
                const double aL = cs.capillary ? DL * sigL * rdx4 : 0.0;
                const double aR = cs.capillary ? DR * sigR * rdx4 : 0.0;
                a[1] = +aL;
                a[2] = -3.0 * aL - aR;
                a[0] = +3.0 * aL + 3.0 * aR;
                a[3] = -aL - 3.0 * aR;
                a[4] = +aR;
                const double DB = face_mobility(ix, jy, 1), DT = face_mobility(ix, jy + 1, 1);
                const double sigB = sigma(face_surf_conc(ix, jy, 1)), sigT = sigma(face_surf_conc(ix, jy + 1, 1));
                const double bB = cs.capillary ? DB * sigB * rdy4 : 0.0;
                const double bT = cs.capillary ? DT * sigT * rdy4 : 0.0;
                a[5] += +bB;
                a[6] += -3.0 * bB - bT;
                a[0] += +3.0 * bB + 3.0 * bT;
                a[7] += -bB - 3.0 * bT;
                a[8] += +bT;
                const double cL = cs.capillary ? DL * sigL * rdx2dy2 : 0.0;
                const double cR = cs.capillary ? DR * sigR * rdx2dy2 : 0.0;
                const double cB = cs.capillary ? DB * sigB * rdx2dy2 : 0.0;
                const double cT = cs.capillary ? DT * sigT * rdx2dy2 : 0.0;
                a[9] += +cL;
                a[2] += -2.0 * cL;
                a[10] += +cL;
                a[11] += +cR;
                a[3] += -2.0 * cR;
                a[12] += +cR;
                a[6] += -cL - cR;
                a[0] += +2.0 * cL + 2.0 * cR;
                a[7] += -cL - cR;
                a[9] += +cB;
                a[6] += -2.0 * cB;
                a[11] += +cB;
                a[10] += +cT;
                a[7] += -2.0 * cT;
                a[12] += +cT;
                a[2] += -cB - cT;
                a[0] += +2.0 * cB + 2.0 * cT;
                a[3] += -cB - cT;

                const double sPo = SS(ix + 2, jy), spo = SS(ix + 1, jy), smo = SS(ix - 1, jy), sMo = SS(ix - 2, jy);
                const double soo = SS(ix, jy);
                double dhdt = 0.0;
                if (cs.capillary) {
                    const double d3sL = (spo - 3.0 * soo + 3.0 * smo - sMo) * rdx * rdx2;
                    const double d3sR = (sPo - 3.0 * spo + 3.0 * soo - smo) * rdx * rdx2;
                    dhdt -= (DR * sigR * d3sR - DL * sigL * d3sL) * rdx;
                }
                if (cs.vertical_gravity) {
                    const double hsP = HH(ix + 1, jy) + spo, hsO = HH(ix, jy) + soo, hsM = HH(ix - 1, jy) + smo;
                    const double gR = DR * cs.Bo * (hsP - hsO) * rdx;
                    const double gL = DL * cs.Bo * (hsO - hsM) * rdx;
                    dhdt += (gR - gL) * rdx;
                }
                if (cs.inplane_gravity) {
                    const double fR = DR * cs.Pgx, fL = DL * cs.Pgx;
                    dhdt -= (fR - fL) * rdx;
                }
                if (cs.centrifugal) {
                    const double xR = xface(ix + 1) - cs.rot_x, xL = xface(ix) - cs.rot_x;
                    dhdt -= (DR * cs.Sp * xR - DL * cs.Sp * xL) * rdx;
                }
                if (cs.marangoni) dhdt -= (marangoni_flux(ix + 1, jy, 0) - marangoni_flux(ix, jy, 0)) * rdx;
                if (cs.capillary) {
                    const double soP = SS(ix, jy + 2), sop = SS(ix, jy + 1), som = SS(ix, jy - 1), soM = SS(ix, jy - 2);
                    const double d3sB = (sop - 3.0 * soo + 3.0 * som - soM) * rdy * rdy2;
                    const double d3sT = (soP - 3.0 * sop + 3.0 * soo - som) * rdy * rdy2;
                    dhdt -= (DT * sigT * d3sT - DB * sigB * d3sB) * rdy;
                    const double smm = SS(ix - 1, jy - 1), smp = SS(ix - 1, jy + 1);
                    const double spm = SS(ix + 1, jy - 1), spp = SS(ix + 1, jy + 1);
                    const double xL_ = ((som - 2.0 * soo + sop) - (smm - 2.0 * smo + smp)) * rdx * rdy2;
                    const double xR_ = ((spm - 2.0 * spo + spp) - (som - 2.0 * soo + sop)) * rdx * rdy2;
                    dhdt -= (DR * sigR * xR_ - DL * sigL * xL_) * rdx;
                    const double yB_ = ((smo - 2.0 * soo + spo) - (smm - 2.0 * som + spm)) * rdy * rdx2;
                    const double yT_ = ((smp - 2.0 * sop + spp) - (smo - 2.0 * soo + spo)) * rdy * rdx2;
                    dhdt -= (DT * sigT * yT_ - DB * sigB * yB_) * rdy;
                }
                if (cs.vertical_gravity) {
                    const double sop = SS(ix, jy + 1), som = SS(ix, jy - 1);
                    const double hsT = HH(ix, jy + 1) + sop, hsO = HH(ix, jy) + soo, hsB = HH(ix, jy - 1) + som;
                    const double gT = DT * cs.Bo * (hsT - hsO) * rdy;
                    const double gB = DB * cs.Bo * (hsO - hsB) * rdy;
                    dhdt += (gT - gB) * rdy;
                }
                if (cs.inplane_gravity) dhdt -= (DT * cs.Pgy - DB * cs.Pgy) * rdy;
                if (cs.centrifugal) {
                    const double yT = yface(jy + 1) - cs.rot_y, yB = yface(jy) - cs.rot_y;
                    dhdt -= (DT * cs.Sp * yT - DB * cs.Sp * yB) * rdy;
                }
                if (cs.marangoni) dhdt -= (marangoni_flux(ix, jy + 1, 1) - marangoni_flux(ix, jy, 1)) * rdy;
                double Evap = cs.Ev * evap_rate(fH, fC, ix, jy);
                if (Evap > 0.0) {
                    Evap = solvent_limit(Evap, fH[I2(ix, jy)], column_mean(fC, ix, jy), dt);
                    Evap = solvent_limit(Evap, tH[I2(ix, jy)], column_mean(tC, ix, jy), dt);
                }
                dhdt -= Evap;
                colSrc[I2(ix, jy)] = 0.0 - Evap - 0.0;
                Bh[I2(ix, jy)] = -dhdt;
            }
        store_face_flux_parts();
    }

    void store_face_flux_parts() {
        const double rdx = 1.0 / dx, rdx2 = rdx * rdx, rdy = 1.0 / dy, rdy2 = rdy * rdy;
        auto Sf = [&](int i, int j) { return S[I2(i, j)]; };
        auto Hf = [&](int i, int j) { return fH[I2(i, j)]; };
        for (int ix = 2; ix <= Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                const double D = face_mobility(ix, jy, 0);
                double cap = 0.0, expl = 0.0;
                if (cs.capillary) {
                    cap = D * sigma(face_surf_conc(ix, jy, 0));
                    double d3s = (Sf(ix + 1, jy) - 3.0 * Sf(ix, jy) + 3.0 * Sf(ix - 1, jy) - Sf(ix - 2, jy)) * rdx * rdx2;
                    d3s += ((Sf(ix, jy + 1) - 2.0 * Sf(ix, jy) + Sf(ix, jy - 1))
                            - (Sf(ix - 1, jy + 1) - 2.0 * Sf(ix - 1, jy) + Sf(ix - 1, jy - 1))) * rdx * rdy2;
                    expl += cap * d3s;
                }
                if (cs.vertical_gravity)
                    expl -= D * cs.Bo * ((Hf(ix, jy) + Sf(ix, jy)) - (Hf(ix - 1, jy) + Sf(ix - 1, jy))) * rdx;
                if (cs.inplane_gravity) expl += D * cs.Pgx;
                if (cs.centrifugal) expl += D * cs.Sp * (xface(ix) - cs.rot_x);
                if (cs.marangoni) expl += marangoni_flux(ix, jy, 0);
                qCapX[I2(ix, jy)] = cap;
                qExpX[I2(ix, jy)] = expl;
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy <= Ny - 2; jy++) {
                const double D = face_mobility(ix, jy, 1);
                double cap = 0.0, expl = 0.0;
                if (cs.capillary) {
                    cap = D * sigma(face_surf_conc(ix, jy, 1));
                    double d3s = (Sf(ix, jy + 1) - 3.0 * Sf(ix, jy) + 3.0 * Sf(ix, jy - 1) - Sf(ix, jy - 2)) * rdy * rdy2;
                    d3s += ((Sf(ix - 1, jy) - 2.0 * Sf(ix, jy) + Sf(ix + 1, jy))
                            - (Sf(ix - 1, jy - 1) - 2.0 * Sf(ix, jy - 1) + Sf(ix + 1, jy - 1))) * rdy * rdx2;
                    expl += cap * d3s;
                }
                if (cs.vertical_gravity)
                    expl -= D * cs.Bo * ((Hf(ix, jy) + Sf(ix, jy)) - (Hf(ix, jy - 1) + Sf(ix, jy - 1))) * rdy;
                if (cs.inplane_gravity) expl += D * cs.Pgy;
                if (cs.centrifugal) expl += D * cs.Sp * (yface(jy) - cs.rot_y);
                if (cs.marangoni) expl += marangoni_flux(ix, jy, 1);
                qCapY[I2(ix, jy)] = cap;
                qExpY[I2(ix, jy)] = expl;
            }
    }

    /* the film's CRS: IA = I + theta dt A on the real rows, the ghost rows
     * (1, -1 to the row they copy) */
    Csr h_system(vector<double> &rhs, const vector<double> &Ax_old) {
        Csr M; M.n = n2; M.ptr.assign(n2 + 1, 0); M.bc.assign(n2, 0);
        rhs.assign(n2, 0.0);
        for (int r = 0; r < n2; ++r) {
            const int ix = r % Nx, jy = r / Nx;
            const int s = h_src(ix, jy);
            if (s >= 0) {
                M.bc[r] = 1;
                M.col.push_back(r); M.val.push_back(1.0);
                M.col.push_back(s); M.val.push_back(-1.0);
            } else {
                const double *a = &Ah[(size_t)r * 13];
                M.col.push_back(r); M.val.push_back(1.0 + theta * dt * a[0]);
                for (int k = 1; k < 13; ++k) {
                    M.col.push_back(I2(ix + hox[k], jy + hoy[k]));
                    M.val.push_back(theta * dt * a[k]);
                }
                rhs[r] = h_old[r] - dt * (1.0 - theta) * Ax_old[r] - dt * Bh[r];
            }
            M.ptr[r + 1] = (int)M.col.size();
        }
        return M;
    }
    /* A phi in storage order; the ghost rows: phi_g - phi_src */
    void h_matvec(const vector<double> &x, vector<double> &y) const {
        y.assign(n2, 0.0);
        for (int r = 0; r < n2; ++r) {
            const int ix = r % Nx, jy = r / Nx;
            const int s = h_src(ix, jy);
            if (s >= 0) { y[r] = 1.0 * x[r] + (-1.0) * x[s]; continue; }
            const double *a = &Ah[(size_t)r * 13];
            double sum = 0.0;
            sum += a[0] * x[r];
            for (int k = 1; k < 13; ++k) sum += a[k] * x[I2(ix + hox[k], jy + hoy[k])];
            y[r] = sum;
        }
    }

    /* SolveStage of the film (:1949-1963) */
    int h_stage(int &iter, double &resid) {
        h_set_matrix();
        vector<double> Ax_old, rhs;
        h_matvec(h_old, Ax_old);
        Csr M = h_system(rhs, Ax_old);
        int it = 0; double rs = 0;
        const int ch = linsolve(M, rhs, h_new, YZ_PC_AMG, 1, cs.lin_maxit, cs.h_eps, it, rs);
        iter += it; resid = rs;
        st_[0] += 1; st_[1] += it;
        if (ch != 0) {
            bool finite = true;
            for (double v : h_new) finite = finite && std::isfinite(v);
            (void)finite;
        }
        return ch;
    }

    /* ConservativeUpdate (:1232-1301) */
    void conservative_update() {
        const double rdx = 1.0 / dx, rdx2 = rdx * rdx, rdy = 1.0 / dy, rdy2 = rdy * rdy;
        copy_ghosts(h_new);
        vector<double> hc(n2);
        for (int i = 0; i < n2; ++i) {
            const double a = h_new[i] * theta;
            const double b = h_old[i] * (1.0 - theta);
            hc[i] = a + b;
        }
        auto Hc = [&](int i, int j) { return hc[I2(i, j)]; };
        for (int ix = 2; ix <= Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                double d3h = (Hc(ix + 1, jy) - 3.0 * Hc(ix, jy) + 3.0 * Hc(ix - 1, jy) - Hc(ix - 2, jy)) * rdx * rdx2;
                d3h += ((Hc(ix, jy + 1) - 2.0 * Hc(ix, jy) + Hc(ix, jy - 1))
                        - (Hc(ix - 1, jy + 1) - 2.0 * Hc(ix - 1, jy) + Hc(ix - 1, jy - 1))) * rdx * rdy2;
                qX[I2(ix, jy)] = qCapX[I2(ix, jy)] * d3h + qExpX[I2(ix, jy)];
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy <= Ny - 2; jy++) {
                double d3h = (Hc(ix, jy + 1) - 3.0 * Hc(ix, jy) + 3.0 * Hc(ix, jy - 1) - Hc(ix, jy - 2)) * rdy * rdy2;
                d3h += ((Hc(ix - 1, jy) - 2.0 * Hc(ix, jy) + Hc(ix + 1, jy))
                        - (Hc(ix - 1, jy - 1) - 2.0 * Hc(ix, jy - 1) + Hc(ix + 1, jy - 1))) * rdy * rdx2;
                qY[I2(ix, jy)] = qCapY[I2(ix, jy)] * d3h + qExpY[I2(ix, jy)];
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                double div = (qX[I2(ix + 1, jy)] - qX[I2(ix, jy)]) * rdx;
                div += (qY[I2(ix, jy + 1)] - qY[I2(ix, jy)]) * rdy;
                h_new[I2(ix, jy)] = h_old[I2(ix, jy)] + dt * (colSrc[I2(ix, jy)] - div);
            }
        copy_ghosts(h_new);
    }

    /* CalVelocity (:1381-1749), the in-plane components the
     * concentration reads: x faces ix = 2 .. Nx-2 and y faces jy = 2 ..
     * Ny-2 of the real rows and columns, levels 2 .. Nz-2 */
    void velocity_face(int ix, int jy, int dir, double *u) const {
        /* a: along the face normal, b: across */
        const int ax = dir == 0, ay = dir == 1;
        auto H = [&](int k) { return fH[I2(ix + k * ax, jy + k * ay)]; };
        auto Sx = [&](int k) { return S[I2(ix + k * ax, jy + k * ay)]; };
        auto C = [&](int k, int kz) { return fC[I3(ix + k * ax, jy + k * ay, kz)]; };
        const double ds = dir == 0 ? dx : dy;
        const double dn = dir == 0 ? dy : dx;   /* the other axis */
        const double ds3 = std::pow(ds, 3.0);
        const double hpo = H(1), hoo = H(0), hmo = H(-1), hMo = H(-2);
        const double spo = Sx(1), soo = Sx(0), smo = Sx(-1), sMo = Sx(-2);
        const double hL = 0.5 * (hmo + hoo), sL = 0.5 * (smo + soo);
        const double xL = (dir == 0 ? xface(ix) - cs.rot_x : yface(jy) - cs.rot_y);
        double Cpoo = C(1, Nz - 2), Cooo = C(0, Nz - 2), Cmoo = C(-1, Nz - 2), CMoo = C(-2, Nz - 2);
        double Cpom = C(1, Nz - 3), Coom = C(0, Nz - 3), Cmom = C(-1, Nz - 3), CMom = C(-2, Nz - 3);
        const double dcdx_suf = 0.5 * ((-Cpoo + 27.0 * Cooo - 27.0 * Cmoo + CMoo) + (-Cpom + 27.0 * Coom - 27.0 * Cmom + CMom)) / (24.0 * ds);
        const double dx3hs = ((hpo + spo) - 3.0 * (hoo + soo) + 3.0 * (hmo + smo) - (hMo + sMo)) / ds3;
        const double dxhs = (-(hpo + spo) + 27.0 * (hoo + soo) - 27.0 * (hmo + smo) + (hMo + sMo)) / (24.0 * ds);
        const double sig = sigma(0.25 * (Cooo + Coom + Cmoo + Cmom));
        Cooo = C(0, 2); Cmoo = C(-1, 2); Coom = C(0, 1); Cmom = C(-1, 1);
        const double muB = visc(0.25 * (Cooo + Coom + Cmoo + Cmom));
        double fB1 = 1.0 / muB, fB2 = sL / muB;
        const double dzz = dz * hL;
        double sum1 = 0, sum2 = 0;
        /* the cross term: from the rows (columns) either side */
        const int bx = ay, by = ax;   /* across the face */
        auto HS = [&](int a, int b) {
            const int i = ix + a * ax + b * bx, j = jy + a * ay + b * by;
            return fH[I2(i, j)] + S[I2(i, j)];
        };
        const double dn2 = dn * dn;
        const double fOp = HS(0, 1), fOm = HS(0, -1), fMp = HS(-1, 1), fMm = HS(-1, -1);
        const double dxyy = ((fOp - 2.0 * (hoo + soo) + fOm) - (fMp - 2.0 * (hmo + smo) + fMm)) / (ds * dn2);
        const double kLap = cs.capillary ? sig * (dx3hs + dxyy) : 0.0;
        const double kMa = cs.marangoni ? cs.As * cs.Ma * dcdx_suf : 0.0;
        const double kBo = cs.vertical_gravity ? -cs.Bo * dxhs : 0.0;
        const double kPg = cs.inplane_gravity ? (dir == 0 ? cs.Pgx : cs.Pgy) : 0.0;
        const double kSp = cs.centrifugal ? cs.Sp * xL : 0.0;
        const double hLsL = hL + sL;
        for (int kz = 2; kz <= Nz - 2; kz++) {
            const double zT = sL + dzz * (-1.5 + (double)kz);
            const double zM = (kz == 2) ? sL + 0.25 * dzz : sL + dzz * (-2.0 + (double)kz);
            if (kz > 2) {
                Coom = Cooo; Cmom = Cmoo;
                Cooo = C(0, kz); Cmoo = C(-1, kz);
            }
            const double muM = (kz == 2) ? visc(0.375 * (Cooo + Cmoo) + 0.125 * (Coom + Cmom))
                                         : visc(0.25 * (Cooo + Coom + Cmoo + Cmom));
            const double muT = visc(0.50 * (Cooo + Cmoo));
            const double fT1 = 1.0 / muT, fT2 = zT / muT, fM1 = 1.0 / muM, fM2 = zM / muM;
            const double dzt = (kz == 2) ? 0.5 * dzz : dzz;
            sum1 += (fB1 + fT1 + 4.0 * fM1) * dzt / 6.0;
            sum2 += (fB2 + fT2 + 4.0 * fM2) * dzt / 6.0;
            fB1 = fT1; fB2 = fT2;
            const double PP = hLsL * sum1 - sum2;
            double uu = 0;
            if (cs.capillary) uu += kLap * PP;
            if (cs.marangoni) uu += kMa * sum1;
            if (cs.vertical_gravity) uu += kBo * PP;
            if (cs.inplane_gravity) uu += kPg * PP;
            if (cs.centrifugal) uu += kSp * PP;
            u[kz] = uu;
        }
    }
    void cal_velocity() {
        std::fill(Ux.begin(), Ux.end(), 0.0);
        std::fill(Uy.begin(), Uy.end(), 0.0);
        vector<double> u(Nz, 0.0);
        for (int ix = 2; ix <= Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                velocity_face(ix, jy, 0, u.data());
                const bool w = wall_face(ix, jy, 0);
                for (int kz = 2; kz <= Nz - 2; kz++) Ux[I3(ix, jy, kz)] = w ? 0.0 : u[kz];
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy <= Ny - 2; jy++) {
                velocity_face(ix, jy, 1, u.data());
                const bool w = wall_face(ix, jy, 1);
                for (int kz = 2; kz <= Nz - 2; kz++) Uy[I3(ix, jy, kz)] = w ? 0.0 : u[kz];
            }
    }

    /* BuildVolumeFluxes (CZ_ConcEquation_2D.cpp:144-263) */
    void layer_mean(const double *u, double *um) const {
        for (int kz = 2; kz <= Nz - 3; kz++) {
            if (kz == 2) um[kz] = u[kz] + (u[kz + 1] - 3.0 * u[kz]) / 18.0;
            else um[kz] = u[kz] + (u[kz + 1] - 2.0 * u[kz] + u[kz - 1]) / 24.0;
        }
    }
    void build_volume_fluxes() {
        vector<double> u(Nz, 0.0), um(Nz, 0.0);
        for (int ix = 2; ix <= Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                for (int kz = 2; kz <= Nz - 2; kz++) u[kz] = Ux[I3(ix, jy, kz)];
                layer_mean(u.data(), um.data());
                const double hf = 0.5 * (fH[I2(ix - 1, jy)] + fH[I2(ix, jy)]);
                double sum = 0.0;
                for (int kz = 2; kz < Nz - 2; kz++) {
                    const double q = hf * um[kz];
                    Qx[I3(ix, jy, kz)] = q;
                    sum += q * dz;
                }
                const double d = qX[I2(ix, jy)] - sum;
                for (int kz = 2; kz < Nz - 2; kz++) Qx[I3(ix, jy, kz)] += d;
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy <= Ny - 2; jy++) {
                for (int kz = 2; kz <= Nz - 2; kz++) u[kz] = Uy[I3(ix, jy, kz)];
                layer_mean(u.data(), um.data());
                const double hf = 0.5 * (fH[I2(ix, jy - 1)] + fH[I2(ix, jy)]);
                double sum = 0.0;
                for (int kz = 2; kz < Nz - 2; kz++) {
                    const double q = hf * um[kz];
                    Qy[I3(ix, jy, kz)] = q;
                    sum += q * dz;
                }
                const double d = qY[I2(ix, jy)] - sum;
                for (int kz = 2; kz < Nz - 2; kz++) Qy[I3(ix, jy, kz)] += d;
            }
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                const double ht = dtdh[I2(ix, jy)];
                double w = 0.0;
                W[I3(ix, jy, 2)] = 0.0;
                for (int kz = 2; kz < Nz - 2; kz++) {
                    double div = ht + (Qx[I3(ix + 1, jy, kz)] - Qx[I3(ix, jy, kz)]) / dx;
                    div += (Qy[I3(ix, jy + 1, kz)] - Qy[I3(ix, jy, kz)]) / dy;
                    w -= dz * div;
                    W[I3(ix, jy, kz + 1)] = w;
                }
                W[I3(ix, jy, Nz - 2)] = -(colSrc[I2(ix, jy)] + 0.0);
            }
    }

    /* slot order of a concentration row as stored: the diagonal, then
     * put()'s order (CZ_ConcEquation_2D.cpp:535-556) */
    static constexpr int cox[21] = {0, 0, -1, 0, 1, -2, -1, 1, 2, -1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int coy[21] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -2, -1, 1, 2, -1, 1, -1, 1};
    static constexpr int coz[21] = {0, -2, -1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 2, 0, 0, 0, 0, -1, -1, 1, 1};

    /* AdvWeights (:575-593): an unlimited scheme's weights, the same for
     * both rows; a limited one's frozen at the field f, in the upwind form
     * for the upwind cell and the downwind form for the other (rowpos) */
    void conc_weights(const double f[4], double F, int rowpos, double w[4]) const {
        if (clim) { lim.weights_at(f, F, rowpos, w); return; }
        for (int j = 0; j < 4; ++j) w[j] = c1w[j] + Sign(F) * c2w[j];
    }

    /* SetMatrixA of the concentration (:290-570) */
    void c_set_matrix() {
        build_volume_fluxes();
        std::fill(Ac.begin(), Ac.end(), 0.0);
        std::fill(Bc.begin(), Bc.end(), 0.0);
        std::fill(react.begin(), react.end(), 0.0);
        const double As = cs.As, Df = cs.Df;
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                auto C = [&](int i, int j, int k) { return fC[I3(i, j, k)]; };
                auto Hh = [&](int i, int j) { return fH[I2(i, j)]; };
                auto Ss = [&](int i, int j) { return S[I2(i, j)]; };
                const bool wallXmin = (ix == 2) && wall[0];
                const bool wallXmax = (ix == Nx - 3) && wall[1];
                const bool wallYmin = (jy == 2) && wall[2];
                const bool wallYmax = (jy == Ny - 3) && wall[3];
                const double hoo = Hh(ix, jy);
                const double hpo = Hh(ix + 1, jy), hPo = Hh(ix + 2, jy), hmo = Hh(ix - 1, jy), hMo = Hh(ix - 2, jy);
                const double spo = Ss(ix + 1, jy), sPo = Ss(ix + 2, jy), smo = Ss(ix - 1, jy), sMo = Ss(ix - 2, jy);
                const double soo = Ss(ix, jy);
                const double hL = 0.5 * (hoo + hmo), hR = 0.5 * (hoo + hpo);
                const double hx0 = (hMo - 8.0 * hmo + 8.0 * hpo - hPo) / (12.0 * dx);
                const double sx0 = (sMo - 8.0 * smo + 8.0 * spo - sPo) / (12.0 * dx);
                const double hxL = (hMo - 27.0 * hmo + 27.0 * hoo - hpo) / (24.0 * dx);
                const double sxL = (sMo - 27.0 * smo + 27.0 * soo - spo) / (24.0 * dx);
                const double hxR = (hmo - 27.0 * hoo + 27.0 * hpo - hPo) / (24.0 * dx);
                const double sxR = (smo - 27.0 * soo + 27.0 * spo - sPo) / (24.0 * dx);
                const double hoP = Hh(ix, jy + 2), hop = Hh(ix, jy + 1), hom = Hh(ix, jy - 1), hoM = Hh(ix, jy - 2);
                const double soP = Ss(ix, jy + 2), sop = Ss(ix, jy + 1), som = Ss(ix, jy - 1), soM = Ss(ix, jy - 2);
                const double hS = 0.5 * (hoo + hom), hN = 0.5 * (hoo + hop);
                const double hy0 = (hoM - 8.0 * hom + 8.0 * hop - hoP) / (12.0 * dy);
                const double sy0 = (soM - 8.0 * som + 8.0 * sop - soP) / (12.0 * dy);
                const double hyS = (hoM - 27.0 * hom + 27.0 * hoo - hop) / (24.0 * dy);
                const double syS = (soM - 27.0 * som + 27.0 * soo - sop) / (24.0 * dy);
                const double hyN = (hom - 27.0 * hoo + 27.0 * hop - hoP) / (24.0 * dy);
                const double syN = (som - 27.0 * soo + 27.0 * sop - soP) / (24.0 * dy);
                for (int kz = 2; kz < Nz - 2; kz++) {
                    const int irow = I3(ix, jy, kz);
                    const double zeta = dz * (-1.5 + (double)kz);
                    const bool bottom = (kz == 2), top = (kz == Nz - 3);
                    double xz[5][5] = {{0}}, yz[5][3] = {{0}};
                    double rct = 0.0;
                    auto X = [&](int di, int dk, double v) { xz[di + 2][dk + 2] += v; };
                    auto Y = [&](int dj, int dk, double v) { if (dj == 0) xz[2][dk + 2] += v; else yz[dj + 2][dk + 1] += v; };
                    {
                        const double QL = Qx[I3(ix, jy, kz)], QR = Qx[I3(ix + 1, jy, kz)];
                        const double fL[4] = {C(ix - 2, jy, kz), C(ix - 1, jy, kz), C(ix, jy, kz), C(ix + 1, jy, kz)};
                        const double fR[4] = {C(ix - 1, jy, kz), C(ix, jy, kz), C(ix + 1, jy, kz), C(ix + 2, jy, kz)};
                        double wL[4], wR[4];
                        conc_weights(fL, QL, 2, wL); conc_weights(fR, QR, 1, wR);
                        for (int t = 0; t < 4; t++) {
                            X(t - 2, 0, +QL * wL[t] / dx);
                            X(t - 1, 0, -QR * wR[t] / dx);
                        }
                    }
                    {
                        const double QS = Qy[I3(ix, jy, kz)], QN = Qy[I3(ix, jy + 1, kz)];
                        const double fS[4] = {C(ix, jy - 2, kz), C(ix, jy - 1, kz), C(ix, jy, kz), C(ix, jy + 1, kz)};
                        const double fN[4] = {C(ix, jy - 1, kz), C(ix, jy, kz), C(ix, jy + 1, kz), C(ix, jy + 2, kz)};
                        double wS[4], wN[4];
                        conc_weights(fS, QS, 2, wS); conc_weights(fN, QN, 1, wN);
                        for (int t = 0; t < 4; t++) {
                            Y(t - 2, 0, +QS * wS[t] / dy);
                            Y(t - 1, 0, -QN * wN[t] / dy);
                        }
                    }
                    {
                        const double WB = W[I3(ix, jy, kz)], WT = W[I3(ix, jy, kz + 1)];
                        if (!bottom) {
                            const double fB[4] = {C(ix, jy, kz - 2), C(ix, jy, kz - 1), C(ix, jy, kz), C(ix, jy, kz + 1)};
                            double wB[4]; conc_weights(fB, WB, 2, wB);
                            for (int t = 0; t < 4; t++) X(0, t - 2, +WB * wB[t] / dz);
                        }
                        if (!top) {
                            const double fT[4] = {C(ix, jy, kz - 1), C(ix, jy, kz), C(ix, jy, kz + 1), C(ix, jy, kz + 2)};
                            double wT[4]; conc_weights(fT, WT, 1, wT);
                            for (int t = 0; t < 4; t++) X(0, t - 1, -WT * wT[t] / dz);
                        } else {
                            rct += WT / dz;
                        }
                    }
                    const double Coo = C(ix, jy, kz);
                    {
                        const double DL = Df * diffu(0.5 * (C(ix - 1, jy, kz) + Coo));
                        const double DR = Df * diffu(0.5 * (Coo + C(ix + 1, jy, kz)));
                        const double AL = (hxL * zeta + sxL) / hL;
                        const double AR = (hxR * zeta + sxR) / hR;
                        const double aL = wallXmin ? 0.0 : hL * DL / dx, aR = wallXmax ? 0.0 : hR * DR / dx;
                        X(-1, 0, +aL / dx);
                        X(0, 0, -aL / dx);
                        X(-1, 1, +aL * AL / (4.0 * dz)); X(0, 1, +aL * AL / (4.0 * dz));
                        X(-1, -1, -aL * AL / (4.0 * dz)); X(0, -1, -aL * AL / (4.0 * dz));
                        X(1, 0, +aR / dx);
                        X(0, 0, -aR / dx);
                        X(0, 1, -aR * AR / (4.0 * dz)); X(1, 1, -aR * AR / (4.0 * dz));
                        X(0, -1, +aR * AR / (4.0 * dz)); X(1, -1, +aR * AR / (4.0 * dz));
                    }
                    {
                        const double DS = Df * diffu(0.5 * (C(ix, jy - 1, kz) + Coo));
                        const double DN = Df * diffu(0.5 * (Coo + C(ix, jy + 1, kz)));
                        const double AS = (hyS * zeta + syS) / hS;
                        const double AN = (hyN * zeta + syN) / hN;
                        const double aS = wallYmin ? 0.0 : hS * DS / dy, aN = wallYmax ? 0.0 : hN * DN / dy;
                        Y(-1, 0, +aS / dy);
                        Y(0, 0, -aS / dy);
                        Y(-1, 1, +aS * AS / (4.0 * dz)); Y(0, 1, +aS * AS / (4.0 * dz));
                        Y(-1, -1, -aS * AS / (4.0 * dz)); Y(0, -1, -aS * AS / (4.0 * dz));
                        Y(1, 0, +aN / dy);
                        Y(0, 0, -aN / dy);
                        Y(0, 1, -aN * AN / (4.0 * dz)); Y(1, 1, -aN * AN / (4.0 * dz));
                        Y(0, -1, +aN * AN / (4.0 * dz)); Y(1, -1, +aN * AN / (4.0 * dz));
                    }
                    for (int side = 0; side < 2; side++) {
                        if (side == 0 && bottom) continue;
                        if (side == 1 && top) continue;
                        const int k0 = (side == 0) ? kz - 1 : kz;
                        const double zf = zeta + (side == 0 ? -0.5 : 0.5) * dz;
                        const double Dz = Df * diffu(0.5 * (C(ix, jy, k0) + C(ix, jy, k0 + 1)));
                        const double hA = hx0 * zf + sx0;
                        const double hAy = hy0 * zf + sy0;
                        const double g = (side == 1) ? +1.0 : -1.0;
                        const double kzz = g * Dz * (As / hoo + (hA * hA + hAy * hAy) / hoo) / (dz * dz);
                        X(0, k0 + 1 - kz, +kzz);
                        X(0, k0 - kz, -kzz);
                        const double kx = -g * Dz * hA / (4.0 * dx) / dz;
                        X(1, k0 - kz, +kx); X(1, k0 + 1 - kz, +kx);
                        X(-1, k0 - kz, -kx); X(-1, k0 + 1 - kz, -kx);
                        const double ky = -g * Dz * hAy / (4.0 * dy) / dz;
                        Y(1, k0 - kz, +ky); Y(1, k0 + 1 - kz, +ky);
                        Y(-1, k0 - kz, -ky); Y(-1, k0 + 1 - kz, -ky);
                    }
                    const double rh = 1.0 / hoo;
                    double offsum = 0.0;
                    double *a = &Ac[(size_t)irow * 21];
                    int slot = 1;
                    auto put = [&](double v) { const double e = -v * rh; a[slot++] = e; offsum += e; };
                    put(xz[2][0]); put(xz[1][1]); put(xz[2][1]); put(xz[3][1]);
                    put(xz[0][2]); put(xz[1][2]); put(xz[3][2]); put(xz[4][2]);
                    put(xz[1][3]); put(xz[2][3]); put(xz[3][3]); put(xz[2][4]);
                    put(yz[0][1]); put(yz[1][1]); put(yz[3][1]); put(yz[4][1]);
                    put(yz[1][0]); put(yz[3][0]); put(yz[1][2]); put(yz[3][2]);
                    react[irow] = -rct * rh;
                    a[0] = -offsum + react[irow];
                    Bc[irow] = 0.0;
                }
            }
        c_set_bc(false);
    }

    /* the row a ghost row of the concentration copies, and its kind: 0
     * interior, 1 a copy across x or y (-1 to src), 2 a z ghost of a real
     * column (-1 to src, right-hand side cRhs) */
    int c_src(int ix, int jy, int kz, int *kind) const {
        *kind = 0;
        if (ix < 2 || ix > Nx - 3) { *kind = 1; return I3(fold(ix, Nx, !wall[0], ix < 2), jy, kz); }
        if (jy < 2 || jy > Ny - 3) { *kind = 1; return I3(ix, fold(jy, Ny, !wall[2], jy < 2), kz); }
        if (kz == 0) { *kind = 2; return I3(ix, jy, 3); }
        if (kz == 1) { *kind = 2; return I3(ix, jy, 2); }
        if (kz == Nz - 2) { *kind = 2; return I3(ix, jy, Nz - 3); }
        if (kz == Nz - 1) { *kind = 2; return I3(ix, jy, Nz - 4); }
        return -1;
    }

    /* SetBoundaryCondition (:588-732): the right-hand sides of the z ghost
     * rows of every real column */
    void c_set_bc(bool atEnd) {
        const double As = cs.As, Df = cs.Df;
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                auto C = [&](int i, int j, int k) { return fC[I3(i, j, k)]; };
                auto Hh = [&](int i, int j) { return fH[I2(i, j)]; };
                auto Ss = [&](int i, int j) { return S[I2(i, j)]; };
                const double hh = Hh(ix, jy);
                const double hx0 = (Hh(ix - 2, jy) - 8.0 * Hh(ix - 1, jy) + 8.0 * Hh(ix + 1, jy) - Hh(ix + 2, jy)) / (12.0 * dx);
                const double sx0 = (Ss(ix - 2, jy) - 8.0 * Ss(ix - 1, jy) + 8.0 * Ss(ix + 1, jy) - Ss(ix + 2, jy)) / (12.0 * dx);
                const double hy0 = (Hh(ix, jy - 2) - 8.0 * Hh(ix, jy - 1) + 8.0 * Hh(ix, jy + 1) - Hh(ix, jy + 2)) / (12.0 * dy);
                const double sy0 = (Ss(ix, jy - 2) - 8.0 * Ss(ix, jy - 1) + 8.0 * Ss(ix, jy + 1) - Ss(ix, jy + 2)) / (12.0 * dy);
                {
                    const int k = 2;
                    const double cb = 0.5 * (C(ix, jy, k - 1) + C(ix, jy, k));
                    const double D = Df * diffu(cb);
                    const double cx = (C(ix + 1, jy, k) - C(ix - 1, jy, k)) / (2.0 * dx);
                    const double cy = (C(ix, jy + 1, k) - C(ix, jy - 1, k)) / (2.0 * dy);
                    const double hA = sx0, hAy = sy0;
                    const double cz = hh * (0.0 / D + hA * cx + hAy * cy) / (As + hA * hA + hAy * hAy);
                    cRhs[I3(ix, jy, 0)] = -3.0 * dz * cz;
                    cRhs[I3(ix, jy, 1)] = -dz * cz;
                }
                {
                    const int k = Nz - 3;
                    const double csf = 0.5 * (C(ix, jy, k) + C(ix, jy, k + 1));
                    const double D = Df * diffu(csf);
                    const double cx = (C(ix + 1, jy, k) - C(ix - 1, jy, k)) / (2.0 * dx);
                    const double cy = (C(ix, jy + 1, k) - C(ix, jy - 1, k)) / (2.0 * dy);
                    const double hA = hx0 + sx0, hAy = hy0 + sy0;
                    double WT = W[I3(ix, jy, Nz - 2)];
                    if (atEnd) {
                        WT = cs.Ev * evap_rate(fH, fC, ix, jy);
                        double cm = 0.0;
                        for (int kk = 2; kk < Nz - 2; kk++) cm += C(ix, jy, kk);
                        cm /= (double)(Nz - 4);
                        WT = solvent_limit(WT, hh, cm, dt);
                    }
                    const double G = WT * csf;
                    const double cz = hh * (G / D + hA * cx + hAy * cy) / (As + hA * hA + hAy * hAy);
                    const double grad = dz * cz;
                    cRhs[I3(ix, jy, Nz - 2)] = grad;
                    cRhs[I3(ix, jy, Nz - 1)] = 3.0 * grad;
                }
            }
    }

    /* ApplyBoundaryRows (:985-1005): four sweeps in row order */
    void apply_boundary_rows(vector<double> &x) const {
        for (int sweep = 0; sweep < 4; sweep++)
            for (int r = 0; r < n3; r++) {
                const int kz = r / (Nx * Ny), jy = (r / Nx) % Ny, ix = r % Nx;
                int kind;
                const int s = c_src(ix, jy, kz, &kind);
                if (s < 0) continue;
                double sum = kind == 2 ? cRhs[r] : 0.0;
                sum -= (-1.0) * x[s];
                x[r] = sum / 1.0;
            }
    }
    /* ExplicitPart (:895-913) */
    void explicit_part(const vector<double> &x, vector<double> &y) const {
        y.assign(n3, 0.0);
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++)
                for (int kz = 2; kz < Nz - 2; kz++) {
                    const int r = I3(ix, jy, kz);
                    const double *a = &Ac[(size_t)r * 21];
                    const double xr = x[r];
                    double s = react[r] * xr;
                    for (int k = 1; k < 21; ++k)
                        s += a[k] * (x[I3(ix + cox[k], jy + coy[k], kz + coz[k])] - xr);
                    y[r] = s;
                }
    }

    /* SolveStage of the concentration (:755-893) */
    int c_stage(int &iter, double &resid) {
        c_set_matrix();
        apply_boundary_rows(c_old);
        vector<double> Ax_old;
        explicit_part(c_old, Ax_old);
        Csr M; M.n = n3; M.ptr.assign(n3 + 1, 0); M.bc.assign(n3, 0);
        vector<double> rhs(n3, 0.0), d(n3, 0.0);
        bool any = false;
        for (int r = 0; r < n3; ++r) {
            const int kz = r / (Nx * Ny), jy = (r / Nx) % Ny, ix = r % Nx;
            int kind;
            const int s = c_src(ix, jy, kz, &kind);
            if (s >= 0) {
                M.bc[r] = 1;
                M.col.push_back(r); M.val.push_back(1.0);
                M.col.push_back(s); M.val.push_back(-1.0);
            } else {
                const double *a = &Ac[(size_t)r * 21];
                M.col.push_back(r); M.val.push_back(1.0 + theta * dt * a[0]);
                for (int k = 1; k < 21; ++k) {
                    M.col.push_back(I3(ix + cox[k], jy + coy[k], kz + coz[k]));
                    M.val.push_back(theta * dt * a[k]);
                }
                rhs[r] = -dt * (Ax_old[r] + Bc[r]);
                if (rhs[r] != 0.0) any = true;
            }
            M.ptr[r + 1] = (int)M.col.size();
        }
        int ch = 0, it = 0;
        double rs = 0;
        if (any) {
            ch = linsolve(M, rhs, d, YZ_PC_AMG, 1, cs.lin_maxit, cs.c_eps, it, rs);
            st_[2] += 1; st_[3] += it;
        }
        iter += it; resid = rs;
        if (ch != 0) {
            bool finite = true;
            for (double v : d) finite = finite && std::isfinite(v);
            if (!finite) std::fill(d.begin(), d.end(), 0.0);
        }
        if (any && ch == 0) {
            {
                vector<double> cn(n3);
                for (int i = 0; i < n3; ++i) cn[i] = c_old[i] + d[i];
                apply_boundary_rows(cn);
                for (int i = 0; i < n3; ++i) d[i] = cn[i] - c_old[i];
            }
            vector<double> ct(n3), Acv;
            for (int i = 0; i < n3; ++i) { const double t = d[i] * theta; ct[i] = t + c_old[i]; }
            explicit_part(ct, Acv);
            double shr = 0.0, sh = 0.0;
            for (int jy = 2; jy < Ny - 2; jy++)
                for (int ix = 2; ix < Nx - 2; ix++) {
                    const double h = fH[I2(ix, jy)];
                    for (int kz = 2; kz < Nz - 2; kz++) {
                        const int r = I3(ix, jy, kz);
                        shr += h * (-dt * (Acv[r] + Bc[r]) - d[r]);
                        sh += h * (1.0 + theta * dt * react[r]);
                    }
                }
            const double alpha = (sh > 0.0) ? shr / sh : 0.0;
            for (int r = 0; r < n3; r++) d[r] += alpha;
        }
        for (int i = 0; i < n3; ++i) c_new[i] = c_old[i] + d[i];
        apply_boundary_rows(c_new);
        return ch;
    }
    /* SettleLimiter (:945-973): a limited scheme's step solved again with
     * the limiter frozen at the midpoint of what came out, until that
     * settles to 1e-12 of the step's change (8 passes at most) */
    int settle_limiter(int &iter, double &resid) {
        if (!clim) return 0;
        vector<double> mid(n3);
        for (int pass = 0; pass < 8; pass++) {
            for (int i = 0; i < n3; ++i) mid[i] = (c_old[i] + c_new[i]) * 0.5;
            const vector<double> prev = c_new;
            fC = mid;
            int it = 0; double rs = 0;
            const int ch = c_stage(it, rs);
            iter += it; resid = rs;
            ++settle_passes;
            if (ch != 0) return ch;
            double dmax = 0, cmax = 0;
            for (int i = 0; i < n3; ++i) {
                dmax = std::max(dmax, std::fabs(c_new[i] - prev[i]));
                cmax = std::max(cmax, std::fabs(c_new[i] - c_old[i]));
            }
            if (!(dmax > 1e-12 * cmax)) break;
        }
        return 0;
    }
    /* SolveAtMidpoint (:935-944) */
    int c_solve_at_mid(const vector<double> &cmid, int &iter, double &resid) {
        fC = cmid;
        int ch = c_stage(iter, resid);
        if (ch == 0) ch = settle_limiter(iter, resid);
        fC = c_old;
        return ch;
    }
    /* Solve (:915-933): predictor, then the midpoint */
    int c_solve(int &iter, double &resid) {
        int ch = c_stage(iter, resid);
        if (ch != 0) return ch;
        vector<double> mid(n3);
        for (int i = 0; i < n3; ++i) mid[i] = (c_old[i] + c_new[i]) * 0.5;
        fC = mid;
        ch = c_stage(iter, resid);
        if (ch == 0) ch = settle_limiter(iter, resid);
        fC = c_old;
        return ch;
    }

    /* RedistributeExcess (:403-441) */
    double redistribute_excess(vector<double> &c) const {
        const int kb = 2, kt = Nz - 3;
        double moved = 0.0;
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                bool any = false;
                for (int kz = kb; kz <= kt; kz++) if (c[I3(ix, jy, kz)] > 1.0) { any = true; break; }
                if (!any) continue;
                for (int kz = kt; kz > kb; kz--) {
                    double &ck = c[I3(ix, jy, kz)];
                    if (ck > 1.0) { const double e = ck - 1.0; ck = 1.0; c[I3(ix, jy, kz - 1)] += e; moved += e; }
                }
                double &cb = c[I3(ix, jy, kb)];
                if (cb > 1.0) {
                    double e = cb - 1.0;
                    cb = 1.0;
                    for (int kz = kb + 1; kz <= kt && e > 0.0; kz++) {
                        double &ck = c[I3(ix, jy, kz)];
                        const double take = std::min(e, 1.0 - ck);
                        if (take > 0.0) { ck += take; e -= take; moved += take; }
                    }
                    if (e > 0.0)
                        for (int kz = kb; kz <= kt; kz++) c[I3(ix, jy, kz)] += e / (double)(kt - kb + 1);
                }
            }
        return moved;
    }

    /* RedistributeDeficit (:449-570) */
    void redistribute_deficit(vector<double> &c, double &moved, double &added) const {
        moved = added = 0.0;
        const int kb = 2, kt = Nz - 3;
        const double nLay = (double)(kt - kb + 1);
        auto H = [&](int ix, int jy) { return fH[I2(ix, jy)]; };
        bool any = false;
        for (int ix = 2; ix < Nx - 2 && !any; ix++)
            for (int jy = 2; jy < Ny - 2 && !any; jy++)
                for (int kz = kb; kz <= kt; kz++) if (c[I3(ix, jy, kz)] < 0.0) { any = true; break; }
        if (!any) return;
        bool columns = false;
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                double def = 0.0, pos = 0.0;
                for (int kz = kb; kz <= kt; kz++) {
                    const double ck = c[I3(ix, jy, kz)];
                    if (ck < 0.0) def -= ck; else pos += ck;
                }
                if (def == 0.0) continue;
                moved += def * H(ix, jy) / nLay;
                if (pos >= def) {
                    const double f = def / pos;
                    for (int kz = kb; kz <= kt; kz++) {
                        double &ck = c[I3(ix, jy, kz)];
                        ck = (ck < 0.0) ? 0.0 : ck - f * ck;
                    }
                } else {
                    const double m = (pos - def) / nLay;
                    for (int kz = kb; kz <= kt; kz++) c[I3(ix, jy, kz)] = m;
                    columns = true;
                }
            }
        if (!columns) return;
        auto colMean = [&](int ix, int jy) {
            double s = 0.0;
            for (int kz = kb; kz <= kt; kz++) s += c[I3(ix, jy, kz)];
            return s / nLay;
        };
        auto scale = [&](int ix, int jy, double g) { for (int kz = kb; kz <= kt; kz++) c[I3(ix, jy, kz)] *= g; };
        auto real = [&](int ix, int jy) { return ix >= 2 && ix < Nx - 2 && jy >= 2 && jy < Ny - 2; };
        const int dxs[4] = {-1, 1, 0, 0}, dys[4] = {0, 0, -1, 1};
        vector<double> room((size_t)n2, 0.0), take((size_t)n2, 0.0);
        bool left = true;
        for (int pass = 0; pass < 4 && left; ++pass) {
            for (int ix = 2; ix < Nx - 2; ix++)
                for (int jy = 2; jy < Ny - 2; jy++) {
                    const double m = colMean(ix, jy), h = H(ix, jy);
                    room[I2(ix, jy)] = (m > 0.0 && h > 0.0) ? h * m : 0.0;
                    take[I2(ix, jy)] = 0.0;
                }
            left = false;
            for (int ix = 2; ix < Nx - 2; ix++)
                for (int jy = 2; jy < Ny - 2; jy++) {
                    const double m = colMean(ix, jy), h = H(ix, jy);
                    if (!(m < 0.0)) continue;
                    if (!(h > 0.0)) { added -= h * m; scale(ix, jy, 0.0); continue; }
                    const double e = -h * m;
                    double r = 0.0;
                    for (int q = 0; q < 4; q++)
                        if (real(ix + dxs[q], jy + dys[q])) r += room[I2(ix + dxs[q], jy + dys[q])];
                    if (!(r > 0.0)) { left = true; continue; }
                    const double f = (e < r) ? e / r : 1.0;
                    for (int q = 0; q < 4; q++)
                        if (real(ix + dxs[q], jy + dys[q])) take[I2(ix + dxs[q], jy + dys[q])] += f;
                    const double mNew = (f < 1.0) ? 0.0 : -(e - r) / h;
                    for (int kz = kb; kz <= kt; kz++) c[I3(ix, jy, kz)] = mNew;
                    if (f >= 1.0) left = true;
                }
            for (int ix = 2; ix < Nx - 2; ix++)
                for (int jy = 2; jy < Ny - 2; jy++) {
                    const double t = take[I2(ix, jy)];
                    if (t > 0.0) {
                        scale(ix, jy, 1.0 - t);
                        if (colMean(ix, jy) < 0.0) left = true;
                    }
                }
        }
        if (!left) return;
        double lack = 0.0, total = 0.0;
        for (int ix = 2; ix < Nx - 2; ix++)
            for (int jy = 2; jy < Ny - 2; jy++) {
                const double m = colMean(ix, jy), h = H(ix, jy);
                if (m < 0.0) { lack -= h * m; scale(ix, jy, 0.0); }
                else if (h > 0.0) total += h * m;
            }
        const double f = (total > 0.0) ? std::min(1.0, lack / total) : 0.0;
        if (f > 0.0)
            for (int ix = 2; ix < Nx - 2; ix++)
                for (int jy = 2; jy < Ny - 2; jy++)
                    if (H(ix, jy) > 0.0 && colMean(ix, jy) > 0.0) scale(ix, jy, 1.0 - f);
        if (total < lack) added += lack - total;
    }

    /* ---- set-up ---- */
    int init(const yz_cz_case &c, const double *h, const double *s, const double *conc, double dt0) {
        cs = c;
        Nx = c.nx + 4; Ny = c.ny + 4; Nz = c.nz + 4;
        n2 = Nx * Ny; n3 = n2 * Nz;
        const double dx0 = (c.xmax - c.xmin) / (double)c.nx, dy0 = (c.ymax - c.ymin) / (double)c.ny;
        xlo = c.xmin - 2 * dx0; ylo = c.ymin - 2 * dy0;
        dx = ((c.xmax + 2 * dx0) - xlo) / (double)Nx;
        dy = ((c.ymax + 2 * dy0) - ylo) / (double)Ny;
        const double dz0 = (1.0 - 0.0) / (double)c.nz;
        dz = ((1.0 + 2 * dz0) - (0.0 - 2 * dz0)) / (double)Nz;
        for (int k = 0; k < 4; ++k) wall[k] = c.bc[k] != YZ_BC_PERIODIC;
        clim = c.conc_scheme == YZ_CONV_KOREN || c.conc_scheme == YZ_CONV_VANLEER;
        lim.vanleer = c.conc_scheme == YZ_CONV_VANLEER;
        {
            double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
            switch (c.conc_scheme) {
            case YZ_CONV_UPWIND1: a[1] = 1; a[2] = 1; b[1] = 1; b[2] = -1; break;
            case YZ_CONV_CENTRAL2: a[1] = 1; a[2] = 1; break;
            case YZ_CONV_CENTRAL4: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; break;
            case YZ_CONV_UPWIND2: a[0] = -1; a[1] = 3; a[2] = 3; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
            case YZ_CONV_QUICK: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
            case YZ_CONV_KK: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -3; b[1] = 9; b[2] = -9; b[3] = 3; break;
            default: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
            }
            double den = 0;
            for (int j = 0; j < 4; ++j) den += a[j];
            for (int j = 0; j < 4; ++j) { c1w[j] = a[j] / den; c2w[j] = b[j] / den; }
        }
        dt = dt0;
        dt_old = dt0;                  /* DT().old starts at the initial step */
        S.assign(s, s + n2);
        h_old.assign(h, h + n2); h_new = h_old;
        fH = h_old; tH = h_old;
        c_old.assign(conc, conc + n3); c_new = c_old;
        fC = c_old; tC = c_old;
        Ah.assign((size_t)n2 * 13, 0.0); Bh.assign(n2, 0.0);
        qCapX.assign(n2, 0.0); qExpX.assign(n2, 0.0); qCapY.assign(n2, 0.0); qExpY.assign(n2, 0.0);
        qX.assign(n2, 0.0); qY.assign(n2, 0.0); colSrc.assign(n2, 0.0); dtdh.assign(n2, 0.0);
        Ux.assign(n3, 0.0); Uy.assign(n3, 0.0); Qx.assign(n3, 0.0); Qy.assign(n3, 0.0); W.assign(n3, 0.0);
        react.assign(n3, 0.0); Ac.assign((size_t)n3 * 21, 0.0); Bc.assign(n3, 0.0); cRhs.assign(n3, 0.0);
        return 0;
    }

    /* ---- the step control (TimeStepTrial, SemiImplicit.cpp:100-148) ---- */
    vector<double> h_old2;             /* _phi_old2: the start of the step before */
    double dt_old = 0;                 /* DT().old */
    int tsc_count = 0;
    /* LocalTruncationCriterion (ImplicitSolver.cpp:271-307) of the film the
     * step's solve gave, over every row */
    double truncation_criterion() const {
        double th = cs.tsc_criterion;
        if (std::fabs(th) <= DBL_MIN) th = 1.0e-2;
        if (std::fabs(dt_old) <= DBL_MIN) return 1.0;
        double hmax = 0;
        for (int i = 0; i < n2; ++i) hmax = std::max(hmax, std::fabs(h_old[i]));
        const double wfloor = cs.tsc_floor * hmax;
        double e = 0;
        for (int i = 0; i < n2; ++i) {
            double w = std::fabs(h_old[i]);
            if (w < wfloor) w = wfloor;
            const double den = (dt_old + dt) * w;
            const double ei = (std::fabs(den) <= DBL_MIN)
                ? 0.0 : 2.0 * (dt / dt_old) * (dt_old * h_new[i] + dt * h_old2[i] - (dt_old + dt) * h_old[i]) / den;
            e = std::max(e, std::fabs(ei));
        }
        return 1.0 - e / th;
    }
    /* TimeAdvance with the step control: the trial repeated at a shorter
     * step until one is accepted (TimeStepAdjustment, ImplicitSolver.cpp
     * :144-190); 1 when the step falls below dt_min. Without it, one trial. */
    int advance(int &lin, int &passes_out, double &resid_out) {
        if (!cs.step_control) return trial(lin, passes_out, resid_out, nullptr);
        h_old2 = h_old;
        int lin_sum = 0;
        for (;;) {
            double crit = -1.0;
            const int rc = trial(lin, passes_out, resid_out, &crit);
            lin_sum += lin;
            if (rc == 0 && crit > 0.0) {
                double dt_next = dt;
                if (crit > 0.90 && dt < cs.dt_max && ++tsc_count >= cs.tsc_keep) {
                    tsc_count = 0;
                    dt_next = std::min(dt / cs.tsc_change, cs.dt_max);
                }
                dt_old = dt;
                dt = dt_next;
                lin = lin_sum;
                return 0;
            }
            ++cuts;
            dt *= cs.tsc_change;
            if (dt < cs.dt_min) { lin = lin_sum; return 1; }
        }
    }
    long long cuts = 0;                /* trials the step control rejected */

    /* ---- one trial: TimeAdvance + Solve + AfterTimeStepTrial; with crit,
     * the film's truncation criterion after its solve, and a film it
     * rejects is not carried on (4, the step's start kept) ---- */
    int trial(int &lin, int &passes_out, double &resid_out, double *crit) {
        lin = 0; passes_out = 0; resid_out = 0;
        const vector<double> h_old_keep = h_old, h_new_keep = h_new;
        h_old = h_new;
        fH = tH; fC = tC;
        int iter = 0; double resid = 0;
        /* Solve (:1988-2020) */
        int ch = h_stage(iter, resid);
        if (ch != 0) { h_old = h_old_keep; h_new = h_new_keep; fH = tH; fC = tC; return 1; }
        vector<double> mid(n2);
        for (int i = 0; i < n2; ++i) mid[i] = (h_old[i] + h_new[i]) * 0.5;
        fH = mid;
        vector<double> cm;
        const bool extrap = ccPrevSet;
        if (extrap) {
            const double r = 0.5 * dt / dtPrev;
            cm = c_new;
            for (int i = 0; i < n3; i++) cm[i] += r * (cm[i] - ccPrev[i]);
            fC = cm;
        }
        ch = h_stage(iter, resid);
        fH = h_old;
        if (extrap) fC = c_new;
        if (ch != 0) { h_old = h_old_keep; h_new = h_new_keep; fH = tH; fC = tC; return 1; }
        if (crit) {
            *crit = truncation_criterion();
            if (!(*crit > 0.0)) { lin = iter; h_old = h_old_keep; h_new = h_new_keep; fH = tH; fC = tC; return 4; }
        }

        /* AfterTimeStepTrial (:207-386) */
        conservative_update();
        fH = h_new;
        int passes = 0;
        const vector<double> c0 = c_new;
        if (!extrap) cm = c0;
        const vector<double> keepOld = c_old, keepNew = c_new;
        const vector<double> ccPrevKeep = ccPrev;
        const double dtPrevKeep = dtPrev;
        const bool ccPrevSetKeep = ccPrevSet;
        ccPrev = c0; dtPrev = dt; ccPrevSet = true;
        const int maxPass = cs.marangoni ? 1 + 8 : 1;
        const double floor = std::max(1e-10, 10.0 * cs.c_eps);
        double e0 = 0, e = 0, omega = 1.0;
        vector<double> r(n3), rPrev;
        int cch = 0, fail = 0;
        int pass;
        for (pass = 0; pass < maxPass; pass++) {
            for (int i = 0; i < n2; ++i) mid[i] = (h_old[i] + h_new[i]) * 0.5;
            if (pass > 0) {
                fH = mid; fC = cm;
                int it2 = 0; double r2 = 0;
                ch = h_stage(it2, r2);
                iter += it2; resid = r2;
                fC = c0;
                if (ch != 0) { fail = 1; break; }
                conservative_update();
                for (int i = 0; i < n2; ++i) mid[i] = (h_old[i] + h_new[i]) * 0.5;
            }
            for (int i = 0; i < n2; ++i) { const double d = h_new[i] - h_old[i]; dtdh[i] = d / dt; }
            fH = mid; fC = cm;
            cal_velocity();
            fC = c0;
            int cit = 0; double crs = 0;
            if (pass > 0) cch = c_solve_at_mid(cm, cit, crs);
            else if (maxPass > 1) { c_old = c_new; cch = c_solve_at_mid(cm, cit, crs); }
            else { c_old = c_new; cch = c_solve(cit, crs); }
            iter += cit;
            if (cch != 0) { fail = 2; break; }
            if (maxPass == 1) break;
            passes = pass;
            for (int i = 0; i < n3; ++i) r[i] = (c0[i] + c_new[i]) * 0.5 - cm[i];
            e = 0;
            for (int i = 0; i < n3; ++i) e = std::max(e, std::fabs(r[i]));
            if (pass == 0) e0 = e;
            const bool settled = !(e > 0.1 * e0) || e <= floor;
            if (pass > 0 && !settled) {
                double num = 0, den = 0;
                for (int i = 0; i < n3; i++) {
                    const double d = r[i] - rPrev[i];
                    num += rPrev[i] * d;
                    den += d * d;
                }
                if (den > 0) omega = -omega * num / den;
            }
            if (settled) break;
            for (int i = 0; i < n3; i++) cm[i] += omega * r[i];
            rPrev = r;
        }
        if (!fail && maxPass > 1 && pass == maxPass) fail = 3;
        fH = h_new;
        lin = iter; resid_out = resid;
        if (fail) {
            c_old = keepOld; c_new = keepNew;
            ccPrev = ccPrevKeep; dtPrev = dtPrevKeep; ccPrevSet = ccPrevSetKeep;
            h_old = h_old_keep; h_new = h_new_keep; fH = tH; fC = tC;
            return fail;
        }
        st_[4] += passes;
        passes_out = passes;
        st_[5] += redistribute_excess(c_new);
        double mv = 0, ad = 0;
        redistribute_deficit(c_new, mv, ad);
        st_[6] += mv; st_[7] += ad;
        fC = c_new;
        c_set_bc(true);
        apply_boundary_rows(c_new);
        fC = c_new;
        tH = fH; tC = fC;
        return 0;
    }
};

constexpr int CzStepper::hox[13];
constexpr int CzStepper::hoy[13];
constexpr int CzStepper::cox[21];
constexpr int CzStepper::coy[21];
constexpr int CzStepper::coz[21];

}  // namespace yzo

extern "C" {

void yzo_cz_case_defaults(yz_cz_case *c) {
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = 16; c->nz = 8;
    c->xmin = 0; c->xmax = 1; c->ymin = 0; c->ymax = 1;
    c->capillary = 1;
    c->As = 1; c->sig_c0 = 0.5; c->visc_c0 = 0.5; c->Df = 1; c->diff_c0 = 0.5;
    c->Ev = 1; c->evap_coef = 1; c->evap_func = YZ_EVAP_RAOULT; c->evap_power = 0.5;
    c->molar_resin = 2200; c->molar_solvent = 106;
    c->conc_scheme = YZ_CONV_UTOPIA;
    c->lin_maxit = 5000; c->h_eps = 1e-13; c->c_eps = 1e-13;
    c->step_control = 0; c->tsc_change = 0.5; c->tsc_criterion = 0.01; c->tsc_floor = 0; c->tsc_keep = 50;
    c->dt_min = 1e-15; c->dt_max = 1e30;
}
void *yzo_cz_create(const yz_cz_case *c, const double *h, const double *s, const double *conc, double dt) {
    auto *st = new yzo::CzStepper();
    st->init(*c, h, s, conc, dt);
    return st;
}
void yzo_cz_destroy(void *p) { delete static_cast<yzo::CzStepper *>(p); }
int yzo_cz_advance(void *p, int *lin, int *passes, double *resid) {
    int l = 0, ps = 0; double r = 0;
    const int rc = static_cast<yzo::CzStepper *>(p)->advance(l, ps, r);
    if (lin) *lin = l;
    if (passes) *passes = ps;
    if (resid) *resid = r;
    return rc;
}
void yzo_cz_get_h(void *p, double *h) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    std::copy(st->h_new.begin(), st->h_new.end(), h);
}
void yzo_cz_get_c(void *p, double *c) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    std::copy(st->c_new.begin(), st->c_new.end(), c);
}
void yzo_cz_volume_fluxes(void *p, double *qx, double *qy, double *w) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    std::copy(st->Qx.begin(), st->Qx.end(), qx);
    std::copy(st->Qy.begin(), st->Qy.end(), qy);
    std::copy(st->W.begin(), st->W.end(), w);
}
/* SettleLimiter's solves over the run (a limited scheme) */
long long yzo_cz_settle(void *p) { return static_cast<yzo::CzStepper *>(p)->settle_passes; }
/* the clock of the step control: dt now and old, the rejected trials */
void yzo_cz_clock(void *p, double *out3) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    out3[0] = st->dt; out3[1] = st->dt_old; out3[2] = (double)st->cuts;
}
void yzo_cz_stats(void *p, double *out8) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    for (int k = 0; k < 8; ++k) out8[k] = st->st_[k];
}
/* the in-plane velocity through the film of the given field (CalVelocity),
 * Ux and Uy of every face the concentration reads */
void yzo_cz_velocity(void *p, const double *h, const double *conc, double *ux, double *uy) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    st->fH.assign(h, h + st->n2);
    st->fC.assign(conc, conc + st->n3);
    st->cal_velocity();
    std::copy(st->Ux.begin(), st->Ux.end(), ux);
    std::copy(st->Uy.begin(), st->Uy.end(), uy);
    st->fH = st->tH; st->fC = st->tC;
}
/* RedistributeExcess / RedistributeDeficit on a caller's concentration
 * with the film h */
double yzo_cz_redistribute(void *p, const double *h, double *conc, double *moved, double *added) {
    auto *st = static_cast<yzo::CzStepper *>(p);
    st->fH.assign(h, h + st->n2);
    std::vector<double> c(conc, conc + st->n3);
    const double ex = st->redistribute_excess(c);
    double mv = 0, ad = 0;
    st->redistribute_deficit(c, mv, ad);
    std::copy(c.begin(), c.end(), conc);
    if (moved) *moved = mv;
    if (added) *added = ad;
    st->fH = st->tH;
    return ex;
}

}  /* extern "C" */
