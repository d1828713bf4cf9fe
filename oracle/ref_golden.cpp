// ref_golden.cpp -- golden-vector generator built against the reference's
// own term kernels (test infrastructure).
//
// Compiled by oracle/build_ref.sh together with the reference's
// Eigen-free translation units (everything under yuzen/src except
// numerics/LinearSolverBackend.cpp, which needs Eigen/AMGCL, and
// mainProg.cpp). It does what the reference's own
// yuzen/tests/test_jacobian.cpp harness does (MakeInput :82-138, a
// Field_WM1D/2D over a MakeGrid grid :140-149, one term's SetMatrixA and
// SetJacobiMatrix :189-269), on smooth non-flat fields, and writes what the
// reference computes:
//   * per case (dimension x term x slip x boundary kind): the CRS of A and
//     of J = A + the term's Jacobian part, row by row in the reference's
//     storage order, and A(h)h (the conservative form for the capillary and
//     vertical-gravity terms, AddConservative; the product otherwise);
//   * the term's FaceFlux through every face of every real cell and its
//     SideRate (what FaceVolumeFluxes sums for the concentration), its
//     face velocities (CalVelocity), and the free energy and potential of
//     EnergyBudget.h with the term's parameters;
//   * Mobility::Face/FaceDK and derivatives over a grid of heights and slip
//     lengths; the Koren and van Leer limiters' face value, weights and
//     derivative; the upwinded terms with either limiter;
//     NonDimensionalNumbers::Dp over contact angles; FaceValue and
//     FaceWeights of the unlimited face schemes, and of the limited ones
//     with the row's cell upwind and downwind;
//   * a restart state file written by the reference's C_RestartState;
//   * the evaporation rate of C_Evaporation and the nozzle's supply of
//     Source_WM2D / Source_WM1D.
// Output: binary files (little-endian) plus a JSON manifest, into the
// directory given as argv[1].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <type_traits>

#include "NonDimensionalNumbers.h"
#include "LubFields.h"
#include "FieldInfo.h"
#include "BoundaryCondition.h"
#include "CRS_Matrix.h"
#include "WellMixedTraits.h"
#include "FiniteDifferences.h"
#include "Mobility.h"
#include "RestartState.h"
#include "Evaporation.h"
#include "TimeSteps.h"
#include "WM2D/Source.h"
#include "WM1D/Source_1D.h"
#include "Output.h"
#include "Restart.h"
#include "EnergyBudget.h"

using namespace yuzen;
using namespace yuzen::grid;

namespace {

const double PI = 3.14159265358979323846;

enum Term { tLaplace, tVertical, tParallel, tCentrifugal, tSoluto, tDisjoining };
const char *kName[] = {"capillary", "vertical_gravity", "inplane_gravity", "centrifugal", "marangoni",
                       "disjoining"};

struct Params {
    double Bo = 0.7, Pgx = 0.4, Pgy = -0.2, omega = 1.140175425099138, Ma = 0.3, theta_deg = 10.0,
           hstar = 0.1, visc_exp = 3.0, slip = 0.0, rotx = 0.13, roty = -0.07;
};

// Everything dimensionless from the start: L0 = h0 = sigma0 = mu0 =
// rho = 1, so that As = 1, Bo = g_z, Pg = g_xy, Sp = omega^2, Ma = the
// surface-tension coefficient (NonDimensionalNumbers.cpp:43-141).
Input MakeInput(Term t, const Params &p) {
    Input inp;
    inp.Physics.ConcField.EnableConvection = inp.Physics.ConcField.EnableDiffusion = false;
    inp.Physics.Source.Enable = false;
    inp.Output.Overwrite = inp.Output.Dimensioned = false;
    inp.Output.FieldData.NeutralFile = inp.Output.FieldData.VTKFile = false;
    inp.Numerics.TimeStepControl.Enable = false;
    inp.Numerics.HeightEquation.LinearSolver.RowScaling = false;
    inp.Numerics.ConcEquation.LinearSolver.RowScaling = false;
    inp.Numerics.ManufacturedSolution.Enable = false;
    inp.Restart.Enable = inp.Restart.ASCII = false;
    inp.LengthScaleType = "custom";
    inp.LengthScale = 1.0;
    inp.Fluid.Density.Value = 1.0;
    inp.Fluid.SurfaceTension.BaseValue = 1.0;
    inp.Fluid.SurfaceTension.BaseConc = 0.5;
    inp.Fluid.SurfaceTension.Coef = p.Ma;
    inp.Fluid.Viscosity.BaseValue = 1.0;
    inp.Fluid.Viscosity.BaseConc = 0.5;
    inp.Fluid.Viscosity.Exponent = p.visc_exp;
    inp.Fluid.Diffusivity.BaseValue = 1.0;
    inp.Fluid.Diffusivity.BaseConc = 0.5;
    inp.Fluid.Diffusivity.Exponent = 0.0;
    inp.Fluid.MolarMass.Solvent = 100.0;
    inp.Fluid.MolarMass.Resin = 2000.0;
    inp.Initial.SurfaceShape.TotalHeight = 1.0;
    inp.Initial.SurfaceShape.PrecursorRatio = p.hstar;
    inp.Physics.ConcField.Model = "None";
    inp.Physics.LaplacePressure.Enable = (t == tLaplace);
    inp.Physics.SolutoCapillarity.Enable = (t == tSoluto);
    inp.Physics.VerticalGravity.Enable = (t == tVertical);
    inp.Physics.VerticalGravity.Acceleration = p.Bo;
    inp.Physics.HorizontalGravity.Enable = (t == tParallel);
    inp.Physics.HorizontalGravity.Acceleration.assign(2, 0.0);
    inp.Physics.HorizontalGravity.Acceleration[0] = p.Pgx;
    inp.Physics.HorizontalGravity.Acceleration[1] = p.Pgy;
    inp.Physics.Centrifugal.Enable = (t == tCentrifugal);
    inp.Physics.Centrifugal.RotationRate = p.omega;
    inp.Physics.Centrifugal.RotationAxis.assign(2, 0.0);
    inp.Physics.Centrifugal.RotationAxis[0] = p.rotx;
    inp.Physics.Centrifugal.RotationAxis[1] = p.roty;
    inp.Physics.Evaporation.Enable = false;
    inp.Physics.Evaporation.Coef = 0.0;
    inp.Physics.DisjoiningPressure.Enable = (t == tDisjoining);
    inp.Physics.DisjoiningPressure.ContactAngle = p.theta_deg;
    inp.Physics.DisjoiningPressure.ExponentN = 3.0;
    inp.Physics.DisjoiningPressure.ExponentM = 2.0;
    inp.Physics.Slip.Enable = p.slip > 0;
    inp.Physics.Slip.Length = p.slip;
    return inp;
}

Input::I_Grid MakeGrid(bool twoD, int nx, int ny) {
    Input::I_Grid g;
    g.Dimension = twoD ? "xyh" : "xh";
    g.Xmin = -0.5; g.Xmax = 0.5;
    g.Ymin = -0.4; g.Ymax = 0.4;
    if (!twoD) { g.Ymin = -1.0; g.Ymax = 1.0; }
    g.Nx = nx; g.Ny = twoD ? ny : 1; g.Nz = 1;
    return g;
}

Input::I_Substrate Flat() { Input::I_Substrate s; s.TotalHeight = 0.0; return s; }

// smooth, non-flat fields over the interior; the ghosts take the value of
// the cell the boundary condition copies (mirror or wrap), as the field
// would hold them in a run
struct Shapes {
    bool periodic;
    double H(double x, double y) const {
        if (periodic) return 0.5 + 0.3 * std::sin(2 * PI * x) * (1.1 + 0.4 * std::cos(2 * PI * y / 0.8));
        return 0.2 + 0.6 * (1 + std::sin(PI * x / 0.5 + 0.3)) * (1.1 + 0.4 * std::cos(PI * y / 0.4 - 0.2));
    }
    double S(double x, double y) const {
        if (periodic) return 0.1 * std::cos(2 * PI * x) * std::sin(2 * PI * y / 0.8);
        return 0.3 * std::cos(0.7 * PI * x / 0.5) * std::sin(0.5 * PI * y / 0.4 + 0.4);
    }
    double C(double x, double y) const {
        if (periodic) return 0.5 + 0.15 * std::sin(2 * PI * x) + 0.1 * std::cos(2 * PI * y / 0.8);
        return 0.5 + 0.15 * std::sin(0.8 * PI * x / 0.5 - 0.5) + 0.1 * std::cos(0.6 * PI * y / 0.4);
    }
};

int fold(int i, int np, bool periodic) {
    const int m = np - 4;
    if (i < 2) return periodic ? i + m : 3 - i;
    if (i > np - 3) return periodic ? i - m : 2 * (np - 3) + 1 - i;
    return i;
}

void write_vec(std::ofstream &f, const std::vector<double> &v) {
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
}
void write_ivec(std::ofstream &f, const std::vector<int> &v) {
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(int));
}

// the energy parameters of the case being dumped (set by run_case): the
// free energy and the potential of EnergyBudget.h on the case's field
PotentialParams g_pp;
bool g_periodic = false;

template <class Tr, class FieldT, class TermT>
void dump_term(TermT &term, FieldT &field, const Shapes &sh, const std::string &path, bool conservative) {
    typedef typename Tr::Grid Grid;
    typedef typename Tr::Site Site;
    Grid *G = field.GetGridPtr();
    int n = 0;
    for (Site s = G->template First<Site>(); s != G->template Last<Site>(); ++s)
        n = std::max(n, Tr::Index(G, s) + 1);
    int nxp = G->nx();
    int nyp = n / nxp;
    const bool twoD = nyp > 1;
    std::vector<double> hv(n), sv(n), cv(n), xs(n), ys(n);
    std::vector<char> inside(n, 0);
    for (Site s = G->template First<Site>(); s != G->template Last<Site>(); ++s) {
        double x, y;
        Tr::SiteXY(G, s, x, y);
        const int i = Tr::Index(G, s);
        xs[i] = x; ys[i] = y;
        inside[i] = (G->CheckGhost(s) == Grid::Inside);
    }
    for (int i = 0; i < n; ++i) {
        const int ix = i % nxp, iy = i / nxp;
        const int fx = fold(ix, nxp, sh.periodic), fy = twoD ? fold(iy, nyp, sh.periodic) : iy;
        const int src = fy * nxp + fx;
        hv[i] = sh.H(xs[src], ys[src]);
        sv[i] = sh.S(xs[src], ys[src]);
        cv[i] = sh.C(xs[src], ys[src]);
    }
    for (Site s = G->template First<Site>(); s != G->template Last<Site>(); ++s) {
        const int i = Tr::Index(G, s);
        field.HH(s) = hv[i]; field.SS(s) = sv[i]; field.CC(s) = cv[i];
    }
    CRS_Matrix A;
    A.Build(n);
    for (int i = 0; i < n; ++i) { A(i, i) = 0.0; if (!inside[i]) A.SetAsBC(i); }
    term.SetMatrixA(A, field);
    Vector<double> hx(n), ax(n);
    for (int i = 0; i < n; ++i) hx[i] = hv[i];
    if (conservative) {
        ax = 0;
        term.AddConservative(ax, hx, A, field);
    } else {
        A.MatVec(ax, hx);
    }
    std::vector<int> rp(n + 1, 0), col;
    std::vector<double> av, jv;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < A.NonZero(i); ++k) { col.push_back(A.Jcol(i, k)); av.push_back(A.Element(i, k)); }
        rp[i + 1] = (int)col.size();
    }
    CRS_Matrix J;
    J.Build(n);
    for (int i = 0; i < n; ++i) { J(i, i) = 0.0; if (!inside[i]) J.SetAsBC(i); }
    term.SetMatrixA(J, field);
    term.SetJacobiMatrix(J, field);
    // J in A's column order (J's insertion order may add entries)
    std::vector<int> rpj(n + 1, 0), colj;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < J.NonZero(i); ++k) { colj.push_back(J.Jcol(i, k)); jv.push_back(J.Element(i, k)); }
        rpj[i + 1] = (int)colj.size();
    }
    std::vector<double> axv(n);
    for (int i = 0; i < n; ++i) axv[i] = inside[i] ? ax[i] : 0.0;
    // FaceFlux of every face of every real cell (the flux along +axis
    // through face d, PhysicsTermBase.h:160-188) and SideRate
    const int ndir = twoD ? 4 : 2;
    std::vector<double> ffv((size_t)n * ndir, 0.0), side(n, 0.0);
    std::vector<Site> svec(Tr::nStencil);
    for (Site s = G->template First<Site>(); s != G->template Last<Site>(); ++s) {
        const int i = Tr::Index(G, s);
        if (!inside[i]) continue;
        G->SetSite(svec, s);
        for (int d = 0; d < ndir; ++d) ffv[(size_t)i * ndir + d] = term.FaceFlux(s, d, field, svec);
        side[i] = term.SideRate(s, field);
    }
    // the term's face velocities (its CalVelocity, summed into the mixed
    // field by Cal_MixedVelocity; the other terms' fields stay zero), before
    // the equation's SetVelocityBC: x on the Xm face of every cell, y on its
    // Ym face, in the cell's row
    std::vector<double> vx(n, 0.0), vy(n, 0.0);
    term.CalVelocity(field);
    field.UU().Cal_MixedVelocity();
    for (int i = 0; i < n; ++i) {
        const int ix = i % nxp, iy = i / nxp;
        if constexpr (std::is_same<FieldT, Field_WM2D>::value) {
            vx[i] = field.UU().UU().x(G->template At<Yedge>(ix, iy));
            vy[i] = field.UU().UU().y(G->template At<Xedge>(ix, iy));
        } else {
            vx[i] = field.UU().UU().x(G->template At<Node>(ix));
        }
    }
    // FreeEnergy and Potential (EnergyBudget.h) with this term's parameters
    std::vector<double> en(1 + n, 0.0);
    en[0] = energy_budget::FreeEnergy(field, g_pp, g_periodic, g_periodic);
    {
        // phi is indexed by the grid's linear index: to rows
        std::vector<double> phi;
        energy_budget::Potential(field, g_pp, phi);
        for (int ii = G->imin(); ii <= G->imax(); ++ii) {
            Site st = G->template Of<Site>(ii);
            const int row = Tr::Index(G, st);
            if (row >= 0 && row < n && ii < (int)phi.size()) en[1 + row] = phi[(size_t)ii];
        }
    }
    std::ofstream f(path, std::ios::binary);
    const int hdr[4] = {n, (int)col.size(), (int)colj.size(), nxp};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_ivec(f, rp); write_ivec(f, col); write_vec(f, av);
    write_ivec(f, rpj); write_ivec(f, colj); write_vec(f, jv);
    write_vec(f, hv); write_vec(f, sv); write_vec(f, cv); write_vec(f, axv);
    write_vec(f, ffv); write_vec(f, side);
    write_vec(f, vx); write_vec(f, vy);
    write_vec(f, en);
}

template <class K>
void build_term(K &k, Input &inp, typename K::Grid *G, NonDimensionalNumbers *nd, typename K::Boundary *bc) {
    k.Build(inp.Physics, G, nd, bc);
}

// the height schemes by their number in the case (YZ_HSCHEME_*): Koren,
// van Leer, then the fixed-coefficient ones
const C_Interpolation_base4::t_Scheme kHeightScheme[] = {
    C_Interpolation_base4::Koren,       C_Interpolation_base4::VanLeer,    C_Interpolation_base4::Upwind_1st,
    C_Interpolation_base4::Central_2nd, C_Interpolation_base4::Central_4th, C_Interpolation_base4::Upwind_2nd,
    C_Interpolation_base4::Upwind_3rd,  C_Interpolation_base4::QUICK,      C_Interpolation_base4::KK};

std::string run_case(const std::string &dir, bool twoD, Term t, double slip, bool periodic, std::ostream &man,
                     bool &first, int scheme = 0) {
    Params p;
    p.slip = slip;
    if (t == tDisjoining) p.hstar = 0.1;
    Input inp = MakeInput(t, p);
    NonDimensionalNumbers nd;
    nd.Build(inp, inp.Initial.SurfaceShape.TotalHeight);
    const int nx = twoD ? 18 : 28, ny = 14;
    Input::I_Grid g = MakeGrid(twoD, nx, ny);
    Shapes sh{periodic};
    const auto bk = periodic ? BoundaryCondition::Periodic : BoundaryCondition::Symmetry;
    std::ostringstream name;
    name << (twoD ? "2d" : "1d") << "_" << kName[t] << "_slip" << (slip > 0 ? 1 : 0) << (periodic ? "_per" : "_sym")
         << (scheme == 1 ? "_vl" : "") << (scheme > 1 ? "_hs" + std::to_string(scheme) : "");
    const std::string path = dir + "/" + name.str() + ".bin";
    const bool cons = (t == tLaplace || t == tVertical);
    g_pp = PotentialParams();
    g_pp.capillary = (t == tLaplace);
    g_pp.dp = (t == tDisjoining) ? nd.Dp() : 0.0;
    g_pp.hstar = nd.HStar();
    g_pp.n = inp.Physics.DisjoiningPressure.ExponentN;
    g_pp.m = inp.Physics.DisjoiningPressure.ExponentM;
    g_pp.bo = (t == tVertical) ? nd.Bo() : 0.0;
    g_pp.nd = &nd;
    g_periodic = periodic;
    // the upwinded terms' face interpolation (Numerics.Convection.Height)
    const C_Interpolation_base4::t_Scheme koren = kHeightScheme[scheme];
    if (twoD) {
        Field_WM2D field;
        field.Build(g, Flat(), nd.ExtentScale());
        BoundaryCondition2D bc(field.GetGridPtr()->nx(), field.GetGridPtr()->ny(), bk, bk, bk, bk);
        C_Grid2D *G = field.GetGridPtr();
        switch (t) {
        case tLaplace: { WM2D_Traits::LaplacePressure k; build_term(k, inp, G, &nd, &bc); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        case tVertical: { WM2D_Traits::VerticalGravity k; build_term(k, inp, G, &nd, &bc); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        case tParallel: { WM2D_Traits::ParallelGravity k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        case tCentrifugal: { WM2D_Traits::Centrifugal k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        case tSoluto: { WM2D_Traits::SolutoCapillarity k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        default: { WM2D_Traits::DisjoiningPressure k; build_term(k, inp, G, &nd, &bc); dump_term<WM2D_Grid>(k, field, sh, path, cons); break; }
        }
    } else {
        Field_WM1D field;
        field.Build(g, Flat(), nd.ExtentScale());
        BoundaryCondition1D bc(field.GetGridPtr()->nx(), bk, bk);
        C_Grid1D *G = field.GetGridPtr();
        switch (t) {
        case tLaplace: { WM1D_Traits::LaplacePressure k; build_term(k, inp, G, &nd, &bc); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        case tVertical: { WM1D_Traits::VerticalGravity k; build_term(k, inp, G, &nd, &bc); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        case tParallel: { WM1D_Traits::ParallelGravity k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        case tCentrifugal: { WM1D_Traits::Centrifugal k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        case tSoluto: { WM1D_Traits::SolutoCapillarity k; build_term(k, inp, G, &nd, &bc); k.SetConvectionScheme(koren); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        default: { WM1D_Traits::DisjoiningPressure k; build_term(k, inp, G, &nd, &bc); dump_term<WM1D_Grid>(k, field, sh, path, cons); break; }
        }
    }
    char buf[1024];
    std::snprintf(buf, sizeof(buf),
                  "%s\n  {\"name\": \"%s\", \"dim\": %d, \"nx\": %d, \"ny\": %d, \"term\": \"%s\", \"periodic\": %d, "
                  "\"xmin\": -0.5, \"xmax\": 0.5, \"ymin\": %.17g, \"ymax\": %.17g, "
                  "\"As\": %.17g, \"Ma\": %.17g, \"Bo\": %.17g, \"Pgx\": %.17g, \"Pgy\": %.17g, \"Sp\": %.17g, "
                  "\"Dp\": %.17g, \"hstar\": %.17g, \"slip\": %.17g, \"rot_x\": %.17g, \"rot_y\": %.17g, "
                  "\"visc_c0\": 0.5, \"visc_exp\": %.17g, \"sig_c0\": 0.5, \"height_scheme\": %d}",
                  first ? "" : ",", name.str().c_str(), twoD ? 2 : 1, nx, twoD ? ny : 1, kName[t], periodic ? 1 : 0,
                  twoD ? -0.4 : -1.0, twoD ? 0.4 : 1.0, nd.As(), nd.Ma(), nd.Bo(), nd.Pgx(), nd.Pgy(), nd.Sp(),
                  nd.Dp(), nd.HStar(), slip, p.rotx, p.roty, p.visc_exp, scheme);
    man << buf;
    first = false;
    return path;
}

void dump_mobility(const std::string &dir) {
    // Mobility (Mobility.h) over heights, separations and slip lengths,
    // both branches of FaceDK (series and closed form) included
    std::vector<double> out;
    const double lambdas[] = {0.0, 1e-6, 1e-3, 0.05, 0.5, 3.0};
    const double hs[] = {1e-3, 0.02, 0.3, 1.0, 7.0};
    const double rel[] = {0.0, 1e-5, 9e-4, 1.1e-3, 0.05, 0.7, 3.0, 100.0};
    for (double l : lambdas) {
        Mobility m;
        m.Set(l);
        for (double h1 : hs)
            for (double r : rel) {
                const double h2 = h1 * (1.0 + r);
                out.push_back(l); out.push_back(h1); out.push_back(h2);
                out.push_back(m.Face(h1, h2)); out.push_back(m.Face_d1(h1, h2)); out.push_back(m.Face_d1(h2, h1));
                out.push_back(m.D(h1)); out.push_back(m.dD(h1)); out.push_back(m.DperH(h1));
            }
    }
    std::ofstream f(dir + "/mobility.bin", std::ios::binary);
    const int hdr[2] = {(int)out.size() / 9, 9};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(f, out);
}

void dump_limiter(const std::string &dir, bool vanleer) {
    // Koren's (or van Leer's) limited face value, downwind weights and
    // derivative (FiniteDifferences.cpp) over four-point stencils through
    // every branch
    C_Interpolation_base4 k;
    k.Build(vanleer ? C_Interpolation_base4::VanLeer : C_Interpolation_base4::Koren);
    std::vector<double> out;
    unsigned long long z = 12345;
    auto rnd = [&]() {
        z = z * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((z >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (int t = 0; t < 400; ++t) {
        double f[4];
        for (int j = 0; j < 4; ++j) f[j] = rnd();
        if (t % 7 == 0) f[2] = f[1];           // a zero difference
        if (t % 11 == 0) f[1] = 0.5 * (f[0] + f[2]);
        const double sign = (t % 2 == 0) ? 1.0 : -1.0;
        double w[4], g[4];
        const double v = k.FaceValue(f, sign);
        k.FaceWeights(f, sign, -1, w);
        const double vd = k.FaceDerivative(f, sign, g);
        for (int j = 0; j < 4; ++j) out.push_back(f[j]);
        out.push_back(sign); out.push_back(v);
        for (int j = 0; j < 4; ++j) out.push_back(w[j]);
        out.push_back(vd);
        for (int j = 0; j < 4; ++j) out.push_back(g[j]);
    }
    std::ofstream fo(dir + (vanleer ? "/limiter_vl.bin" : "/limiter.bin"), std::ios::binary);
    const int hdr[2] = {(int)out.size() / 15, 15};
    fo.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(fo, out);
}

void dump_limited_weights(const std::string &dir) {
    // the limited schemes' FaceValue and FaceWeights with the row's cell at
    // L (1) and at R (2) (FiniteDifferences.cpp:153-190: the upwind and the
    // downwind form), over four-point stencils with zero differences and
    // both flow signs
    std::vector<double> out;
    unsigned long long z = 4242;
    auto rnd = [&]() {
        z = z * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((z >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (int vl = 0; vl < 2; ++vl) {
        C_Interpolation_base4 k;
        k.Build(vl ? C_Interpolation_base4::VanLeer : C_Interpolation_base4::Koren);
        for (int t = 0; t < 240; ++t) {
            double f[4], w[4];
            for (int j = 0; j < 4; ++j) f[j] = rnd();
            if (t % 5 == 0) f[1] = f[0];            // a zero upwind difference
            if (t % 7 == 0) f[2] = f[1];            // a zero downwind difference
            if (t % 13 == 0) f[3] = f[2];
            if (t % 11 == 0) f[1] = 0.5 * (f[0] + f[2]);
            const double sign = (t % 2 == 0) ? 1.0 : -1.0;
            const int rowpos = (t / 2) % 2 == 0 ? 1 : 2;
            const double v = k.FaceValue(f, sign);
            k.FaceWeights(f, sign, rowpos, w);
            out.push_back(vl); out.push_back(sign); out.push_back(rowpos);
            for (int j = 0; j < 4; ++j) out.push_back(f[j]);
            out.push_back(v);
            for (int j = 0; j < 4; ++j) out.push_back(w[j]);
        }
    }
    std::ofstream fo(dir + "/limited_weights.bin", std::ios::binary);
    const int hdr[2] = {(int)out.size() / 12, 12};
    fo.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(fo, out);
}

void dump_schemes(const std::string &dir) {
    // the unlimited face schemes of the concentration's advection
    // (FiniteDifferences.cpp:22-93, 152-170): FaceValue and FaceWeights
    const C_Interpolation_base4::t_Scheme sc[] = {
        C_Interpolation_base4::Upwind_1st, C_Interpolation_base4::Central_2nd, C_Interpolation_base4::Central_4th,
        C_Interpolation_base4::Upwind_2nd, C_Interpolation_base4::Upwind_3rd, C_Interpolation_base4::QUICK,
        C_Interpolation_base4::KK};
    std::vector<double> out;
    unsigned long long z = 777;
    auto rnd = [&]() {
        z = z * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((z >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (int k = 0; k < 7; ++k) {
        C_Interpolation_base4 in;
        in.Build(sc[k]);
        for (int t = 0; t < 20; ++t) {
            double f[4], w[4];
            for (int j = 0; j < 4; ++j) f[j] = rnd();
            const double sign = (t % 3 == 0) ? 0.0 : (t % 3 == 1 ? 0.7 : -1.3);
            const double v = in.FaceValue(f, sign);
            in.FaceWeights(f, sign, -1, w);
            out.push_back(k); out.push_back(sign);
            for (int j = 0; j < 4; ++j) out.push_back(f[j]);
            out.push_back(v);
            for (int j = 0; j < 4; ++j) out.push_back(w[j]);
        }
    }
    std::ofstream fo(dir + "/schemes.bin", std::ios::binary);
    const int hdr[2] = {(int)out.size() / 11, 11};
    fo.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(fo, out);
}

void dump_dp(const std::string &dir) {
    // Dp of NonDimensionalNumbers::Build over contact angle, H*, n, m
    std::vector<double> out;
    const double th[] = {5, 10, 15, 20, 40};
    const double hst[] = {0.01, 0.05, 0.1};
    const double nm[][2] = {{3, 2}, {4, 3}, {9, 3}};
    for (double a : th)
        for (double hstar : hst)
            for (auto &e : nm) {
                Params p;
                p.theta_deg = a; p.hstar = hstar;
                Input inp = MakeInput(tDisjoining, p);
                inp.Physics.DisjoiningPressure.ExponentN = e[0];
                inp.Physics.DisjoiningPressure.ExponentM = e[1];
                NonDimensionalNumbers nd;
                nd.Build(inp, 1.0);
                out.push_back(a); out.push_back(hstar); out.push_back(e[0]); out.push_back(e[1]);
                out.push_back(nd.Dp());
            }
    std::ofstream f(dir + "/dp.bin", std::ios::binary);
    const int hdr[2] = {(int)out.size() / 5, 5};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(f, out);
}

}  // namespace

// a restart state file written by the reference's own C_RestartState
// (RestartState.cpp:43-81), with the keys and order of
// WellMixed_ModelBase::WriteState (WellMixedModelBase.h:896-927) that a
// run's next step depends on, for a 1-D film of 8 cells (12 rows with the
// ghosts; the correction's ghosts are copies of the cells they mirror)
void dump_state(const std::string &dir) {
    const int n = 12;
    std::vector<double> h(n), s(n), c(n), old(n), corr(n);
    for (int i = 0; i < n; ++i) {
        h[i] = 1.0 + 0.1 * std::sin(0.7 * i + 0.3);
        s[i] = 0.01 * i;
        c[i] = 0.5 + 0.01 * std::cos((double)i);
        old[i] = h[i] - 1e-3 * std::cos((double)i);
    }
    for (int i = 2; i < n - 2; ++i) corr[i] = 1e-6 * std::sin((double)i);
    corr[0] = corr[3]; corr[1] = corr[2]; corr[n - 2] = corr[n - 3]; corr[n - 1] = corr[n - 4];
    C_RestartState st;
    st.SetText("model", "Well-Mixed-1D");
    st.Set("time", 0.0375);
    st.Set("step", 41.0);
    std::vector<double> dt(3);
    dt[0] = 1.0e-3; dt[1] = 5.0e-4; dt[2] = 5.0e-4;
    st.Set("dt", dt);
    st.Set("adjustCount.height", 7.0);
    st.SetVector("height.old", old);
    st.SetVector("height.correction", corr);
    st.SetVector("solver.height", std::vector<double>());
    st.SetVector("solver.conc", std::vector<double>());
    st.SetVector("field.h", h);
    st.SetVector("field.s", s);
    st.SetVector("field.c", c);
    st.Write(dir + "/ref_state_wm1d.state");
}

// the evaporation rate of the reference's C_Evaporation (Evaporation.cpp
// :62-143) at one cell of a 2-D field, over laws, precursor ramps, h and c:
// rows [law (0 Raoult, 1 PowerN), PowerN, ConcR, PrecursorRatio, Coef, h,
// c, E]; and the nozzle's supply of the reference's Source_WM2D and
// Source_WM1D (SourceBase.cpp, WM2D/Source.cpp, WM1D/Source_1D.cpp) at every
// row of a grid at several times: rows [dim, t, row, S], with the scales and
// inputs in a header row
void dump_evap_source(const std::string &dir) {
    Params p;
    std::vector<double> ev;
    struct Law { const char *func; double n, concr, hstar, coef; };
    const Law laws[] = {{"Raoult", 0.5, 0.0, 0.0, 1.0}, {"PowerN", 0.5, 0.0, 0.0, 0.7},
                        {"PowerN", 0.0, 0.1, 0.0, 1.3}, {"PowerN", 1.5, 0.05, 0.0, 1.0},
                        {"Raoult", 0.5, 0.0, 0.05, 1.0}};
    const double hs[] = {0.01, 0.05, 0.07, 0.5, 1.0, 2.5};
    const double cs[] = {-0.1, 0.0, 0.3, 0.5, 0.9, 1.0, 1.2};
    for (const Law &L : laws) {
        Input inp = MakeInput(tLaplace, p);
        inp.Physics.Evaporation.Enable = true;
        inp.Physics.Evaporation.Coef = L.coef;
        inp.Physics.Evaporation.ConcFunc = L.func;
        inp.Physics.Evaporation.PowerN = L.n;
        inp.Physics.Evaporation.ConcR = L.concr;
        inp.Physics.Evaporation.SpatialFunc = "uniform";
        inp.Initial.SurfaceShape.PrecursorRatio = L.hstar;
        inp.Grid = MakeGrid(true, 6, 5);
        C_Evaporation e;
        e.Build(inp, 1.0);
        Field_WM2D field;
        field.Build(inp.Grid, Flat(), 1.0);
        C_Grid2D *G = field.GetGridPtr();
        const Zface z = G->At<Zface>(3, 3);
        for (double h : hs)
            for (double c : cs) {
                field.HH(z) = h;
                field.CC(z) = c;
                ev.push_back(std::string(L.func) == "PowerN" ? 1 : 0);
                ev.push_back(L.n); ev.push_back(L.concr); ev.push_back(L.hstar); ev.push_back(L.coef);
                ev.push_back(h); ev.push_back(c); ev.push_back(e(field, z));
            }
    }
    {
        std::ofstream f(dir + "/evap.bin", std::ios::binary);
        const int hdr[2] = {(int)ev.size() / 8, 8};
        f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        write_vec(f, ev);
    }
    // the source: a rotating arm with ramps, 3x3 sub-regions (2-D) / 5 (1-D)
    Input inp = MakeInput(tLaplace, p);
    inp.Physics.Source.Enable = true;
    inp.Physics.Source.Radius = 0.13;
    inp.Physics.Source.InletRate = 0.8;
    inp.Physics.Source.TimeConstant = 0.05;
    inp.Physics.Source.ArmLength = 0.25;
    inp.Physics.Source.ArmCenter.assign(2, 0.0);
    inp.Physics.Source.ArmCenter[0] = 0.02;
    inp.Physics.Source.ArmCenter[1] = -0.03;
    inp.Physics.Source.RotationRate = 3.0;
    inp.Physics.Source.StartAngle = 0.4;
    inp.Physics.Source.RampTime = 0.2;
    inp.Physics.Source.PlateauTime = 0.5;
    std::vector<double> out;
    const double times[] = {0.0, 0.01, 0.15, 0.45, 0.8, 1.0};
    for (int twoD = 1; twoD >= 0; --twoD) {
        inp.Grid = MakeGrid(twoD == 1, 24, 20);
        inp.Physics.Source.SubRegions = twoD ? 3 : 5;
        NonDimensionalNumbers nd;
        nd.Build(inp, inp.Initial.SurfaceShape.TotalHeight);
        C_TimeSteps ts;
        ts.Build(1e-3, 0.0, 1e9, false);
        out.push_back(-1); out.push_back(nd.ExtentScale()); out.push_back(nd.HeightScale()); out.push_back(nd.TimeScale());
        if (twoD) {
            Field_WM2D field;
            field.Build(inp.Grid, Flat(), nd.ExtentScale());
            C_Grid2D *G = field.GetGridPtr();
            Source_WM2D src;
            src.Build(inp.Physics, G, &nd, &ts);
            for (double t : times) {
                src.SetEvalTime(t);
                src.updateFraction();
                for (int iy = 0; iy < G->ny(); ++iy)
                    for (int ix = 0; ix < G->nx(); ++ix) {
                        out.push_back(2); out.push_back(t); out.push_back(iy * G->nx() + ix);
                        out.push_back(src(G->At<Zface>(ix, iy)));
                    }
            }
        } else {
            Field_WM1D field;
            field.Build(inp.Grid, Flat(), nd.ExtentScale());
            C_Grid1D *G = field.GetGridPtr();
            Source_WM1D src;
            src.Build(inp.Physics, G, &nd, &ts);
            for (double t : times) {
                src.SetEvalTime(t);
                src.updateFraction();
                for (int ix = 0; ix < G->nx(); ++ix) {
                    out.push_back(1); out.push_back(t); out.push_back(ix);
                    out.push_back(src(G->At<Xedge>(ix)));
                }
            }
        }
    }
    std::ofstream f(dir + "/source.bin", std::ios::binary);
    const int hdr[2] = {(int)out.size() / 4, 4};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(f, out);
}

// The ASCII field file (.dat) that OutputNeutral writes (Output.cpp:830-920)
// from Field_WM1D_Out / Field_WM2D_Out (LubFields.cpp:204-290, 322-372,
// 602-900): on a smooth film with a velocity field of known values, without
// and with dimensions; then what ReadAscii_WM1D / ReadAscii_WM2D
// (Restart.cpp:236-311) make of the first file with H_coef 1.7; and the
// VTK files OutputVTK writes (Output.cpp:375-637) of the same field with the
// five terms' own velocity fields set. Per dimension: neutral_<d>d_in.bin
// (padded h, s, c, face x and y velocities, the terms' x and y velocities,
// the time and the time scale), neutral_<d>d_{nd,dim}00000.{dat,vts},
// neutral_<d>d_read.bin (padded h, s, c after the read); the _in.bin ends
// with FieldInfo of the field.
template <class FieldT, class OutT>
void neutral_case(const std::string &dir, bool twoD) {
    Params p;
    Input inp = MakeInput(tLaplace, p);
    const int nx = twoD ? 18 : 28, ny = 14;
    inp.Grid = MakeGrid(twoD, nx, ny);
    NonDimensionalNumbers nd;
    nd.Build(inp, inp.Initial.SurfaceShape.TotalHeight);
    C_TimeSteps ts;
    ts.Build(1.23456789e-3, 0.0, 1e9, false);
    ++ts;   // the .vts carries the clock (rounded to 7 digits): one step on
    FieldT field;
    field.Build(inp.Grid, Flat(), nd.ExtentScale());
    auto *G = field.GetGridPtr();
    const int nxp = nx + 4, nyp = twoD ? ny + 4 : 1, n = nxp * nyp;
    std::vector<double> hv(n), sv(n), cv(n), vx(n), vy(n, 0.0);
    // the five terms' own velocity fields the .vts writes beside the mixed
    // one: capillary, Marangoni, vertical gravity, in-plane gravity, spin
    std::vector<std::vector<double>> tx(5, std::vector<double>(n, 0.0)), ty(5, std::vector<double>(n, 0.0));
    Shapes sh{false};
    const double dx = G->dx();
    const double dy = twoD ? (0.8 + 4.0 * 0.8 / ny) / nyp : 0.0;
    for (int iy = 0; iy < nyp; ++iy)
        for (int ix = 0; ix < nxp; ++ix) {
            const int i = iy * nxp + ix;
            const double x = -0.5 - 2.0 * (1.0 / nx) + (ix + 0.5) * dx;
            const double y = twoD ? -0.4 - 2.0 * (0.8 / ny) + (iy + 0.5) * dy : 0.0;
            hv[i] = sh.H(x, y); sv[i] = sh.S(x, y); cv[i] = sh.C(x, y);
            vx[i] = 0.1 * std::sin(0.37 * ix + 0.11 * iy) + 0.05 * iy;
            if (twoD) vy[i] = 0.07 * std::cos(0.23 * ix - 0.41 * iy) - 0.02 * ix;
            for (int k = 0; k < 5; ++k) {
                tx[k][i] = 0.01 * (k + 1) * std::sin(0.29 * ix + 0.17 * (k + 1) * iy + k);
                if (twoD) ty[k][i] = 0.02 * (k + 1) * std::cos(0.31 * ix - 0.13 * iy + 0.5 * k);
            }
        }
    for (int iy = 0; iy < nyp; ++iy)
        for (int ix = 0; ix < nxp; ++ix) {
            const int i = iy * nxp + ix;
            if constexpr (std::is_same<FieldT, Field_WM2D>::value) {
                const Zface z = G->template At<Zface>(ix, iy);
                field.HH(z) = hv[i]; field.SS(z) = sv[i]; field.CC(z) = cv[i];
                field.UU().UU().x(G->template At<Yedge>(ix, iy)) = vx[i];
                field.UU().UU().y(G->template At<Xedge>(ix, iy)) = vy[i];
                auto *U = &field.UU();
                std::remove_reference_t<decltype(U->U_LaplacePressure())> *terms[5] = {&U->U_LaplacePressure(), &U->U_SolutoCapillarity(), &U->U_VerticalGravity(),
                                  &U->U_HorizontalGravity(), &U->U_Centrifugal()};
                for (int k = 0; k < 5; ++k) {
                    terms[k]->x(G->template At<Yedge>(ix, iy)) = tx[k][i];
                    terms[k]->y(G->template At<Xedge>(ix, iy)) = ty[k][i];
                }
            } else {
                const Xedge z = G->template At<Xedge>(ix);
                field.HH(z) = hv[i]; field.SS(z) = sv[i]; field.CC(z) = cv[i];
                field.UU().UU().x(G->template At<Node>(ix)) = vx[i];
                auto *U = &field.UU();
                std::remove_reference_t<decltype(U->U_LaplacePressure())> *terms[5] = {&U->U_LaplacePressure(), &U->U_SolutoCapillarity(), &U->U_VerticalGravity(),
                                  &U->U_HorizontalGravity(), &U->U_Centrifugal()};
                for (int k = 0; k < 5; ++k) terms[k]->x(G->template At<Node>(ix)) = tx[k][i];
            }
        }
    const std::string tag = std::string("neutral_") + (twoD ? "2d" : "1d");
    {
        std::ofstream f(dir + "/" + tag + "_in.bin", std::ios::binary);
        const int hdr[2] = {n, nxp};
        f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        write_vec(f, hv); write_vec(f, sv); write_vec(f, cv); write_vec(f, vx); write_vec(f, vy);
        for (int k = 0; k < 5; ++k) { write_vec(f, tx[k]); write_vec(f, ty[k]); }
        const double clock[2] = {ts.NowTime(), nd.TimeScale()};
        f.write(reinterpret_cast<const char *>(clock), sizeof(clock));
        // FieldInfo of the field (FieldInfo.cpp:52-136), what History.dat
        // prints of it
        FieldInfo fi;
        fi.Visit(field);
        const double info[9] = {fi.FluidVol(), fi.ResinVol(), fi.AveConc(), fi.MaxHS(), fi.MinHS(),
                                fi.MaxH(),     fi.MinH(),     fi.MaxC(),    fi.MinC()};
        f.write(reinterpret_cast<const char *>(info), sizeof(info));
    }
    for (int dimd = 0; dimd < 2; ++dimd) {
        OutT out;
        out.Build(field, 1.3e-3, 2.5e-6, 0.07, dimd == 1);
        out.ConvertFromStg(field);
        OutputNeutral w;
        w.Build(ts, nd, dimd == 1, FilePath(dir), tag + (dimd ? "_dim" : "_nd"), 0);
        w.Write(out);
        OutputVTK v;
        v.Build(ts, nd, dimd == 1, FilePath(dir), tag + (dimd ? "_dim" : "_nd"), 0);
        v.Write(out);
    }
    FieldT back;
    back.Build(inp.Grid, Flat(), nd.ExtentScale());
    Input::I_Restart ri;
    ri.Enable = true; ri.ASCII = true; ri.FileName = dir + "/" + tag + "_nd00000.dat";
    ri.RestartTime = 0; ri.FileIndex = 0;
    bool ok;
    if constexpr (std::is_same<FieldT, Field_WM2D>::value) ok = C_Restart::ReadAscii_WM2D(ri, back, 1.7);
    else ok = C_Restart::ReadAscii_WM1D(ri, back, 1.7);
    std::vector<double> rh(n), rs(n), rc(n);
    auto *B = back.GetGridPtr();
    for (int iy = 0; iy < nyp; ++iy)
        for (int ix = 0; ix < nxp; ++ix) {
            const int i = iy * nxp + ix;
            if constexpr (std::is_same<FieldT, Field_WM2D>::value) {
                const Zface z = B->template At<Zface>(ix, iy);
                rh[i] = back.HH(z); rs[i] = back.SS(z); rc[i] = back.CC(z);
            } else {
                const Xedge z = B->template At<Xedge>(ix);
                rh[i] = back.HH(z); rs[i] = back.SS(z); rc[i] = back.CC(z);
            }
        }
    std::ofstream f(dir + "/" + tag + "_read.bin", std::ios::binary);
    const int hdr[2] = {n, ok ? 1 : 0};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    write_vec(f, rh); write_vec(f, rs); write_vec(f, rc);
}

void dump_neutral(const std::string &dir) {
    neutral_case<Field_WM2D, Field_WM2D_Out>(dir, true);
    neutral_case<Field_WM1D, Field_WM1D_Out>(dir, false);
}

int main(int argc, char **argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: ref_golden OUTDIR\n"); return 1; }
    const std::string dir = argv[1];
    std::ofstream man(dir + "/manifest.json");
    man << "{\"cases\": [";
    bool first = true;
    const Term all[] = {tLaplace, tVertical, tParallel, tCentrifugal, tSoluto, tDisjoining};
    for (int d2 = 0; d2 < 2; ++d2)
        for (Term t : all)
            for (double slip : {0.0, 0.05})
                run_case(dir, d2 == 1, t, slip, false, man, first);
    for (int d2 = 0; d2 < 2; ++d2)
        for (Term t : all) run_case(dir, d2 == 1, t, 0.05, true, man, first);
    // the upwinded terms with van Leer's limiter
    for (int d2 = 0; d2 < 2; ++d2)
        for (Term t : {tParallel, tCentrifugal, tSoluto})
            for (double slip : {0.0, 0.05})
                run_case(dir, d2 == 1, t, slip, false, man, first, 1);
    // UTOPIA, with and without slip, 1-D and 2-D; the other fixed schemes
    // in 2-D
    for (int d2 = 0; d2 < 2; ++d2)
        for (Term t : {tParallel, tCentrifugal, tSoluto})
            for (double slip : {0.0, 0.05})
                run_case(dir, d2 == 1, t, slip, false, man, first, 6);
    for (int sch : {2, 3, 4, 5, 7, 8})
        for (Term t : {tParallel, tCentrifugal, tSoluto})
            run_case(dir, true, t, sch == 5 ? 0.05 : 0.0, false, man, first, sch);
    man << "\n]}\n";
    dump_mobility(dir);
    dump_limiter(dir, false);
    dump_limiter(dir, true);
    dump_dp(dir);
    dump_schemes(dir);
    dump_limited_weights(dir);
    dump_state(dir);
    dump_evap_source(dir);
    dump_neutral(dir);
    std::printf("golden vectors written to %s\n", dir.c_str());
    return 0;
}
