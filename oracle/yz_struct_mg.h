/* yz_struct_mg.h -- test infrastructure, part of the CPU oracle (included by
 * yz_oracle.cpp; nothing in the product path links it).
 *
 * The linear solve of the device path restated on the CPU: the padded CRS
 * system (ghost rows x_g - x_src = r_g) folded onto the interior cells as
 * stencil planes, row-scaled, and solved by BiCGSTAB right-preconditioned
 * with the structured smoothed-aggregation multigrid the device builds
 * (3x3 aggregates, P = (I - omega D^-1 A) P_t with omega = 4/3 / rho,
 * rho the Gershgorin bound of D^-1 A; the Galerkin 5x5 coarse operators;
 * SPAI0 pre- and post-smoothing; the dense inverse of the coarsest level).
 * It is the method of the reference's AMGCL smoothed_aggregation + spai0
 * preconditioner (LinearSolverBackend.cpp:662-679) with the aggregates fixed
 * by the grid instead of found from the matrix's strong connections, which
 * the oracle's Amg (yz_oracle.cpp) restates.
 *
 * Why the oracle carries it: a Newton pass's correction is the error the
 * Eisenstat-Walker-truncated solve of the pass before it left (relative
 * residual up to 1e-4), and two multigrids meet that tolerance with
 * different error vectors. With this mode the oracle's solves take the
 * device's error vectors up to rounding, so the Newton control can be
 * compared pass for pass and the Newton counts asserted equal. Its own
 * correctness is pinned by the tests of the solution (the residual bar of
 * the backend contract), the V-cycle's contraction, the prolongator's
 * constant-preserving rows and equal results on a transposed grid. */
#pragma once

namespace yzo {

namespace smg {

constexpr int COARSE_MAX = 128;
constexpr int MAX_LEVELS = 12;

/* a level: nx x ny interior cells, ns value planes val[k][i]; the column
 * of slot k of row i is the cell at offset (ox, oy) folded into the grid
 * (mirrored at a Symmetry edge: -1 -> 0, -2 -> 1; wrapped when periodic) */
struct Op {
    int dim = 2, nx = 0, ny = 1, N = 0, ns = 0, perx = 0, pery = 0;
    int ox[25] = {0}, oy[25] = {0};
};

inline int fold_axis(int j, int n, int periodic) {
    if (j >= 0 && j < n) return j;
    if (periodic) { j %= n; if (j < 0) j += n; return j; }
    if (j < 0) j = -1 - j;
    if (j >= n) j = 2 * n - 1 - j;
    if (j < 0) j = 0;
    if (j >= n) j = n - 1;
    return j;
}
inline int col_of(const Op &op, int ix, int iy, int k) {
    const int jx = fold_axis(ix + op.ox[k], op.nx, op.perx);
    const int jy = op.dim == 2 ? fold_axis(iy + op.oy[k], op.ny, op.pery) : 0;
    return jy * op.nx + jx;
}

inline Op box(int dim, int nx, int ny, int perx, int pery) {
    Op op;
    op.dim = dim; op.nx = nx; op.ny = dim == 2 ? ny : 1; op.N = op.nx * op.ny;
    op.perx = perx; op.pery = pery;
    if (dim == 2) {
        op.ns = 25;
        for (int q = 0; q < 25; ++q) { op.ox[q] = q % 5 - 2; op.oy[q] = q / 5 - 2; }
    } else {
        op.ns = 5;
        for (int q = 0; q < 5; ++q) { op.ox[q] = q - 2; op.oy[q] = 0; }
    }
    return op;
}

/* y = A x, each row's slots summed in slot order */
inline void apply(const Op &op, const vector<double> &val, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
    for (int i = 0; i < op.N; ++i) {
        const int ix = i % op.nx, iy = i / op.nx;
        double s = 0;
        for (int k = 0; k < op.ns; ++k) s += val[(size_t)k * op.N + i] * x[col_of(op, ix, iy, k)];
        y[i] = s;
    }
}

struct Mat {
    Op op;
    const vector<double> *val = nullptr;
    int n = 0;
    void mv(const double *x, double *y) const { apply(op, *val, x, y); }
};

inline int agg_size(int I, int n) { return std::min(3 * I + 3, n) - 3 * I; }
inline int coarse_diff(int J, int I, int nc, int periodic) {
    int d = J - I;
    if (periodic) {
        if (d > nc / 2) d -= nc;
        if (d < -(nc / 2)) d += nc;
    }
    return d;
}
inline int agg_offset(int i, int J, int n, int periodic) {
    int u = i - 3 * J;
    if (periodic) {
        if (u > 5) u -= n;
        if (u < -3) u += n;
    }
    return u;
}

struct Level {
    Op op;
    vector<double> val, minv, dg, P, R, x, b, r;
    int ncx = 0, ncy = 0;
    double rho = 0;
};

struct Hierarchy {
    vector<Level> lv;
    vector<double> inv;   /* the coarsest level's inverse, row-major */
    int nco = 0;

    /* the levels' shapes: 3x3 aggregates while the level holds more than
     * COARSE_MAX cells and a periodic axis keeps at least 5 aggregates */
    void shapes(const Op &op0) {
        lv.clear();
        lv.emplace_back();
        lv[0].op = op0;
        while ((int)lv.size() < MAX_LEVELS) {
            Level &L = lv.back();
            if (L.op.N <= COARSE_MAX) break;
            const int ncx = (L.op.nx + 2) / 3;
            const int ncy = L.op.dim == 2 ? (L.op.ny + 2) / 3 : 1;
            if (L.op.perx && ncx < 5) break;
            if (L.op.dim == 2 && L.op.pery && ncy < 5) break;
            if (ncx * ncy >= L.op.N) break;
            L.ncx = ncx; L.ncy = ncy;
            Level C;
            C.op = box(L.op.dim, ncx, ncy, L.op.perx, L.op.pery);
            lv.push_back(std::move(C));
        }
    }

    /* SPAI0 m_i = a_ii / sum_j a_ij^2, the diagonal and the Gershgorin bound
     * of D^-1 A, over each row's slots merged by column */
    static void level_diag(Level &L) {
        const Op &op = L.op;
        L.minv.assign(op.N, 0.0);
        L.dg.assign(op.N, 0.0);
        double rho = 0;
#pragma omp parallel for schedule(static) reduction(max : rho) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) {
            const int ix = i % op.nx, iy = i / op.nx;
            int cols[25];
            double vals[25];
            int m = 0;
            for (int k = 0; k < op.ns; ++k) {
                const double v = L.val[(size_t)k * op.N + i];
                const int c = col_of(op, ix, iy, k);
                int q = 0;
                for (; q < m; ++q) if (cols[q] == c) break;
                if (q == m) { cols[m] = c; vals[m] = v; ++m; }
                else vals[q] += v;
            }
            double d = 0, ss = 0, off = 0;
            for (int q = 0; q < m; ++q) {
                ss += vals[q] * vals[q];
                if (cols[q] == i) d = vals[q]; else off += std::fabs(vals[q]);
            }
            L.minv[i] = ss > 0 ? d / ss : 0.0;
            L.dg[i] = d;
            const double g = d != 0 ? (std::fabs(d) + off) / std::fabs(d) : 0.0;
            rho = std::max(rho, g);
        }
        L.rho = rho;
    }

    /* the smoothed prolongator P(i, agg(i) + e), e in [-1,1]^dim, kept per
     * fine cell and, as R, per coarse cell by the fine cell's offset */
    static void build_P(Level &L, int ncN) {
        const Op &op = L.op;
        const bool two = op.dim == 2;
        const int NP = two ? 9 : 3;
        const double omega = L.rho > 0 ? (4.0 / 3.0) / L.rho : 0.0;
        L.P.assign((size_t)NP * op.N, 0.0);
        L.R.assign((size_t)(two ? 81 : 9) * ncN, 0.0);
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) {
            const int ix = i % op.nx, iy = i / op.nx;
            const int Ix = ix / 3, Iy = iy / 3;
            const double w = L.dg[i] != 0 ? omega / L.dg[i] : 0.0;
            int ek[25];
            double vk[25];
            for (int k = 0; k < op.ns; ++k) {
                const int c = col_of(op, ix, iy, k);
                const int jx = c % op.nx, jy = c / op.nx;
                const int Jx = jx / 3, Jy = jy / 3;
                const int ex = coarse_diff(Jx, Ix, L.ncx, op.perx);
                const int ey = two ? coarse_diff(Jy, Iy, L.ncy, op.pery) : 0;
                const double tJ = 1.0 / std::sqrt((double)(agg_size(Jx, op.nx) * (two ? agg_size(Jy, op.ny) : 1)));
                ek[k] = two ? (ey + 1) * 3 + (ex + 1) : ex + 1;
                vk[k] = w * L.val[(size_t)k * op.N + i] * tJ;
            }
            const double tI = 1.0 / std::sqrt((double)(agg_size(Ix, op.nx) * (two ? agg_size(Iy, op.ny) : 1)));
            for (int e = 0; e < NP; ++e) {
                double pv = e == NP / 2 ? tI : 0.0;
                for (int k = 0; k < op.ns; ++k) if (ek[k] == e) pv -= vk[k];
                L.P[(size_t)e * op.N + i] = pv;
                const int ex = two ? e % 3 - 1 : e - 1;
                const int ey = two ? e / 3 - 1 : 0;
                int Jx = Ix + ex, Jy = Iy + ey;
                if (op.perx) Jx = (Jx + L.ncx) % L.ncx;
                else if (Jx < 0 || Jx >= L.ncx) continue;
                if (two) {
                    if (op.pery) Jy = (Jy + L.ncy) % L.ncy;
                    else if (Jy < 0 || Jy >= L.ncy) continue;
                }
                const int u = agg_offset(ix, Jx, op.nx, op.perx) + 3;
                const int v = two ? agg_offset(iy, Jy, op.ny, op.pery) + 3 : 0;
                L.R[(size_t)(v * 9 + u) * ncN + Jy * L.ncx + Jx] = pv;
            }
        }
    }

    /* the Galerkin product P^T A P on the grid of aggregates (5x5 box) */
    static void build_RAP(const Level &L, Level &C) {
        const Op &op = L.op, &cop = C.op;
        const bool two = op.dim == 2;
        const int NF = two ? 25 : 5;
        vector<double> AP((size_t)NF * op.N, 0.0);
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) {
            const int ix = i % op.nx, iy = i / op.nx;
            const int Ix = ix / 3, Iy = iy / 3;
            int dxk[25], dyk[25], ck[25];
            double ak[25];
            for (int k = 0; k < op.ns; ++k) {
                const int c = col_of(op, ix, iy, k);
                ck[k] = c;
                ak[k] = L.val[(size_t)k * op.N + i];
                dxk[k] = coarse_diff((c % op.nx) / 3, Ix, L.ncx, op.perx);
                dyk[k] = two ? coarse_diff((c / op.nx) / 3, Iy, L.ncy, op.pery) : 0;
            }
            for (int f = 0; f < NF; ++f) {
                const int fx = two ? f % 5 - 2 : f - 2;
                const int fy = two ? f / 5 - 2 : 0;
                double acc = 0;
                for (int k = 0; k < op.ns; ++k) {
                    const int ex = fx - dxk[k], ey = fy - dyk[k];
                    if (ex < -1 || ex > 1 || ey < -1 || ey > 1) continue;
                    const int e = two ? (ey + 1) * 3 + (ex + 1) : ex + 1;
                    acc += ak[k] * L.P[(size_t)e * op.N + ck[k]];
                }
                AP[(size_t)f * op.N + i] = acc;
            }
        }
        C.val.assign((size_t)cop.ns * cop.N, 0.0);
#pragma omp parallel for schedule(static) if (cop.N > YZO_PAR_MIN)
        for (int I = 0; I < cop.N; ++I) {
            const int Ix = I % cop.nx, Iy = I / cop.nx;
            double ac[25] = {0};
            for (int ey = two ? -1 : 0; ey <= (two ? 1 : 0); ++ey)
                for (int ex = -1; ex <= 1; ++ex) {
                    int Kx = Ix + ex, Ky = Iy + ey;
                    bool ok = true;
                    if (cop.perx) Kx = (Kx + cop.nx) % cop.nx;
                    else ok = ok && Kx >= 0 && Kx < cop.nx;
                    if (two) {
                        if (cop.pery) Ky = (Ky + cop.ny) % cop.ny;
                        else ok = ok && Ky >= 0 && Ky < cop.ny;
                    }
                    if (!ok) continue;
                    const int pe = two ? (1 - ey) * 3 + (1 - ex) : 1 - ex;
                    const int x0 = 3 * Kx, x1 = std::min(3 * Kx + 3, op.nx);
                    const int y0 = two ? 3 * Ky : 0, y1 = two ? std::min(3 * Ky + 3, op.ny) : 1;
                    for (int fy = y0; fy < y1; ++fy)
                        for (int fx = x0; fx < x1; ++fx) {
                            const int i = fy * op.nx + fx;
                            const double pv = L.P[(size_t)pe * op.N + i];
                            for (int gy = two ? -2 : 0; gy <= (two ? 2 : 0); ++gy)
                                for (int gx = -2; gx <= 2; ++gx) {
                                    const int qx = ex + gx, qy = ey + gy;
                                    if (qx < -2 || qx > 2 || qy < -2 || qy > 2) continue;
                                    const int g = two ? (gy + 2) * 5 + (gx + 2) : gx + 2;
                                    const int q = two ? (qy + 2) * 5 + (qx + 2) : qx + 2;
                                    ac[q] += pv * AP[(size_t)g * op.N + i];
                                }
                        }
                }
            for (int q = 0; q < cop.ns; ++q) C.val[(size_t)q * cop.N + I] = ac[q];
        }
    }

    /* Gauss-Jordan inverse with row pivoting (the first row of the largest
     * |pivot|), the row swaps undone as column swaps */
    static void invert(int n, vector<double> &M) {
        vector<int> perm(n);
        for (int k = 0; k < n; ++k) {
            int pr = k;
            double best = -1;
            for (int r = k; r < n; ++r) {
                const double a = std::fabs(M[(size_t)r * n + k]);
                if (a > best) { best = a; pr = r; }
            }
            perm[k] = pr;
            if (pr != k)
                for (int j = 0; j < n; ++j) std::swap(M[(size_t)k * n + j], M[(size_t)pr * n + j]);
            const double piv = M[(size_t)k * n + k];
            const double pinv = piv != 0 ? 1.0 / piv : 0.0;
            for (int j = 0; j < n; ++j) M[(size_t)k * n + j] = j == k ? pinv : M[(size_t)k * n + j] * pinv;
            for (int r = 0; r < n; ++r) {
                if (r == k) continue;
                for (int j = 0; j < n; ++j) {
                    if (j == k) continue;
                    M[(size_t)r * n + j] -= M[(size_t)r * n + k] * M[(size_t)k * n + j];
                }
            }
            for (int r = 0; r < n; ++r)
                if (r != k) M[(size_t)r * n + k] = -M[(size_t)r * n + k] * M[(size_t)k * n + k];
        }
        for (int k = n - 1; k >= 0; --k) {
            const int pr = perm[k];
            if (pr != k)
                for (int r = 0; r < n; ++r) std::swap(M[(size_t)r * n + k], M[(size_t)r * n + pr]);
        }
    }

    /* the hierarchy of the fine planes val0 */
    void build(const Op &op0, const vector<double> &val0) {
        shapes(op0);
        lv[0].val = val0;
        for (size_t l = 0; l < lv.size(); ++l) {
            Level &L = lv[l];
            level_diag(L);
            L.x.assign(L.op.N, 0.0); L.b.assign(L.op.N, 0.0); L.r.assign(L.op.N, 0.0);
            if (l + 1 == lv.size()) break;
            Level &C = lv[l + 1];
            build_P(L, C.op.N);
            build_RAP(L, C);
        }
        const Level &Lc = lv.back();
        nco = Lc.op.N;
        inv.assign((size_t)nco * nco, 0.0);
        for (int i = 0; i < nco; ++i) {
            const int ix = i % Lc.op.nx, iy = i / Lc.op.nx;
            for (int k = 0; k < Lc.op.ns; ++k)
                inv[(size_t)i * nco + col_of(Lc.op, ix, iy, k)] += Lc.val[(size_t)k * nco + i];
        }
        invert(nco, inv);
    }

    void dense(const double *b, double *x) const {
        for (int i = 0; i < nco; ++i) {
            double s = 0;
            for (int j = 0; j < nco; ++j) s += inv[(size_t)i * nco + j] * b[j];
            x[i] = s;
        }
    }

    /* r_c = R r: each coarse cell sums the weights of the fine cells
     * (3J - 3 .. 3J + 5 per axis) against their residual */
    static void restrict_to(const Level &L, const Level &C, const vector<double> &r, vector<double> &bc) {
        const Op &op = L.op, &cop = C.op;
        const bool two = op.dim == 2;
#pragma omp parallel for schedule(static) if (cop.N > YZO_PAR_MIN)
        for (int I = 0; I < cop.N; ++I) {
            const int Ix = I % cop.nx, Iy = I / cop.nx;
            double s = 0;
            for (int v = 0; v < (two ? 9 : 1); ++v)
                for (int u = 0; u < 9; ++u) {
                    int gx = 3 * Ix - 3 + u, gy = two ? 3 * Iy - 3 + v : 0;
                    bool ok = true;
                    if (op.perx) gx = ((gx % op.nx) + op.nx) % op.nx;
                    else ok = ok && gx >= 0 && gx < op.nx;
                    if (two) {
                        if (op.pery) gy = ((gy % op.ny) + op.ny) % op.ny;
                        else ok = ok && gy >= 0 && gy < op.ny;
                    }
                    const double rv = ok ? r[(size_t)gy * op.nx + gx] : 0.0;
                    s += L.R[(size_t)(v * 9 + u) * cop.N + I] * rv;
                }
            bc[I] = s;
        }
    }

    /* x += P x_c */
    static void prolong(const Level &L, const Level &C, vector<double> &x) {
        const Op &op = L.op, &cop = C.op;
        const bool two = op.dim == 2;
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) {
            const int ix = i % op.nx, iy = i / op.nx;
            const int Ix = ix / 3, Iy = iy / 3;
            double s = 0;
            for (int ey = two ? -1 : 0; ey <= (two ? 1 : 0); ++ey) {
                int Jy = Iy + ey;
                if (two) {
                    if (cop.pery) Jy = (Jy + cop.ny) % cop.ny;
                    else if (Jy < 0 || Jy >= cop.ny) continue;
                }
                for (int ex = -1; ex <= 1; ++ex) {
                    int Jx = Ix + ex;
                    if (cop.perx) Jx = (Jx + cop.nx) % cop.nx;
                    else if (Jx < 0 || Jx >= cop.nx) continue;
                    const int e = two ? (ey + 1) * 3 + (ex + 1) : ex + 1;
                    s += L.P[(size_t)e * op.N + i] * C.x[(size_t)Jy * cop.nx + Jx];
                }
            }
            x[i] += s;
        }
    }

    /* one V-cycle, out = M^-1 in: SPAI0 from zero, the residual restricted,
     * the coarse correction prolonged, SPAI0 once more */
    void cycle(size_t l, const vector<double> &b, vector<double> &out) {
        Level &L = lv[l];
        const Op &op = L.op;
        if (l + 1 == lv.size()) { dense(b.data(), out.data()); return; }
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) L.x[i] = L.minv[i] * b[i];
        apply(op, L.val, L.x.data(), L.r.data());
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) L.r[i] = b[i] - L.r[i];
        Level &C = lv[l + 1];
        restrict_to(L, C, L.r, C.b);
        cycle(l + 1, C.b, C.x);
        prolong(L, C, L.x);
        apply(op, L.val, L.x.data(), L.r.data());
#pragma omp parallel for schedule(static) if (op.N > YZO_PAR_MIN)
        for (int i = 0; i < op.N; ++i) out[i] = L.x[i] + L.minv[i] * (b[i] - L.r[i]);
    }
    void apply_pc(const vector<double> &in, vector<double> &out) {
        out.resize(in.size());
        cycle(0, in, out);
    }
};

struct Pc {
    Hierarchy *h = nullptr;
    void apply(const vector<double> &b, vector<double> &x) { h->apply_pc(b, x); }
};

/* damped Jacobi, 0.72 / (the slots that fold onto the diagonal) */
struct JacobiPc {
    vector<double> dinv;
    void build(const Op &op, const vector<double> &val) {
        dinv.assign(op.N, 0.0);
        for (int i = 0; i < op.N; ++i) {
            const int ix = i % op.nx, iy = i / op.nx;
            double d = 0;
            for (int k = 0; k < op.ns; ++k)
                if (col_of(op, ix, iy, k) == i) d += val[(size_t)k * op.N + i];
            dinv[i] = d != 0 ? 0.72 / d : 0.0;
        }
    }
    void apply(const vector<double> &b, vector<double> &x) const {
        x.resize(b.size());
        for (size_t i = 0; i < b.size(); ++i) x[i] = dinv[i] * b[i];
    }
};

}  // namespace smg

}  // namespace yzo
