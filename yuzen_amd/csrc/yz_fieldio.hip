// yz_fieldio.hip -- the field files of a well-mixed film: the ASCII file
// (.dat) of the reference's output (OutputNeutral::Write, Output.cpp:830-920,
// of Field_WM1D_Out / Field_WM2D_Out, LubFields.cpp:204-290, 322-372,
// 602-900), its ASCII restart, which reads such a file as an initial
// condition (C_Restart::ReadAscii_WM1D / ReadAscii_WM2D, Restart.cpp:
// 170-311), and the VTK structured grid (.vts, OutputVTK::Write,
// Output.cpp:375-637).
//
// The file holds nxp-2 (x nyp-2) points: the wall faces and the real cell
// centres. A point on the edge takes the mean of the ghost and the first real
// cell (four cells at a corner); its velocity is the face velocity of the
// wall face, a point inside takes the mean of its two faces. Numbers are
// written with 17 digits after the point, so a double survives the file
// exactly. Host code only.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/yuzen_gpu.h"

namespace yz {
void set_error(const std::string &s);
}

namespace {

struct Layout {
    int dim, nxp, nyp, nox, noy;   // padded extents, points in the file
    double xmin, xmax, ymin, ymax, dx, dy;
};

bool layout_of(const yz_case *c, Layout &L) {
    if (!c || (c->dim != 1 && c->dim != 2) || c->nx < 3 || (c->dim == 2 && c->ny < 3)) return false;
    L.dim = c->dim;
    L.nxp = c->nx + 4;
    L.nyp = c->dim == 2 ? c->ny + 4 : 1;
    L.nox = L.nxp - 2;
    L.noy = c->dim == 2 ? L.nyp - 2 : 1;
    L.xmin = c->xmin; L.xmax = c->xmax; L.ymin = c->ymin; L.ymax = c->ymax;
    // the grid's spacing over the padded extent (Grids.cpp:88-110)
    const double dx0 = (c->xmax - c->xmin) / (double)c->nx;
    const double xlo = c->xmin - 2 * dx0;
    L.dx = ((c->xmax + 2 * dx0) - xlo) / L.nxp;
    if (c->dim == 2) {
        const double dy0 = (c->ymax - c->ymin) / (double)c->ny;
        const double ylo = c->ymin - 2 * dy0;
        L.dy = ((c->ymax + 2 * dy0) - ylo) / L.nyp;
    } else {
        L.dy = 1;
    }
    return true;
}

// the points' coordinates: the wall, the first centre half a cell in, then
// cell by cell, the last centre and the far wall
std::vector<double> coords(int n, double lo, double hi, double d) {
    std::vector<double> v(n, 0.0);
    v[0] = lo;
    if (n != 1) {
        v[1] = lo + 0.5 * d;
        v[n - 1] = hi;
        v[n - 2] = hi - 0.5 * d;
        for (int i = 2; i < n - 2; ++i) v[i] = v[i - 1] + d;
    }
    return v;
}

// a padded cell field at the file's points
std::vector<double> points_of(const Layout &L, const double *v) {
    std::vector<double> out((size_t)L.nox * L.noy, 0.0);
    const int X = L.nxp, Y = L.nyp;
    if (L.dim == 1) {
        for (int ix = 1; ix <= X - 2; ++ix) {
            double r;
            if (ix == 1) r = 0.5 * (v[ix] + v[ix + 1]);
            else if (ix == X - 2) r = 0.5 * (v[ix] + v[ix - 1]);
            else r = v[ix];
            out[ix - 1] = r;
        }
        return out;
    }
    auto at = [&](int ix, int iy) { return v[(size_t)iy * X + ix]; };
    for (int iy = 1; iy <= Y - 2; ++iy)
        for (int ix = 1; ix <= X - 2; ++ix) {
            const bool xl = ix == 1, xh = ix == X - 2, yl = iy == 1, yh = iy == Y - 2;
            const double o = at(ix, iy);
            double r;
            if (xl && yl) r = 0.25 * (o + at(ix + 1, iy) + at(ix, iy + 1) + at(ix + 1, iy + 1));
            else if (xh && yl) r = 0.25 * (o + at(ix - 1, iy) + at(ix, iy + 1) + at(ix - 1, iy + 1));
            else if (xl && yh) r = 0.25 * (o + at(ix + 1, iy) + at(ix, iy - 1) + at(ix + 1, iy - 1));
            else if (xh && yh) r = 0.25 * (o + at(ix - 1, iy) + at(ix, iy - 1) + at(ix - 1, iy - 1));
            else if (xl) r = 0.5 * (o + at(ix + 1, iy));
            else if (xh) r = 0.5 * (o + at(ix - 1, iy));
            else if (yl) r = 0.5 * (o + at(ix, iy + 1));
            else if (yh) r = 0.5 * (o + at(ix, iy - 1));
            else r = o;
            out[(size_t)(iy - 1) * L.nox + ix - 1] = r;
        }
    return out;
}

// the face velocities (x on the Xm face, y on the Ym face of each padded
// cell) at the file's points
void velocity_points(const Layout &L, const double *vx, const double *vy, std::vector<double> &ux,
                     std::vector<double> &uy) {
    const int X = L.nxp, Y = L.nyp;
    ux.assign((size_t)L.nox * L.noy, 0.0);
    uy.assign((size_t)L.nox * L.noy, 0.0);
    if (L.dim == 1) {
        for (int ix = 1; ix <= X - 2; ++ix) {
            double r;
            if (ix == 1) r = vx[ix + 1];
            else if (ix == X - 2) r = vx[ix];
            else r = 0.5 * (vx[ix + 1] + vx[ix]);
            ux[ix - 1] = r;
        }
        return;
    }
    auto fx = [&](int ix, int iy) { return vx[(size_t)iy * X + ix]; };
    auto fy = [&](int ix, int iy) { return vy[(size_t)iy * X + ix]; };
    for (int iy = 1; iy <= Y - 2; ++iy)
        for (int ix = 1; ix <= X - 2; ++ix) {
            const bool xl = ix == 1, xh = ix == X - 2, yl = iy == 1, yh = iy == Y - 2;
            double a, b;
            if (xl && yl) {
                a = 0.5 * (fx(ix + 1, iy) + fx(ix + 1, iy + 1));
                b = 0.5 * (fy(ix, iy + 1) + fy(ix + 1, iy + 1));
            } else if (xh && yl) {
                a = 0.5 * (fx(ix, iy) + fx(ix, iy + 1));
                b = 0.5 * (fy(ix, iy + 1) + fy(ix - 1, iy + 1));
            } else if (xl && yh) {
                a = 0.5 * (fx(ix + 1, iy) + fx(ix + 1, iy - 1));
                b = 0.5 * (fy(ix, iy) + fy(ix + 1, iy));
            } else if (xh && yh) {
                a = 0.5 * (fx(ix, iy) + fx(ix, iy - 1));
                b = 0.5 * (fy(ix, iy) + fy(ix - 1, iy));
            } else if (xl) {
                a = fx(ix + 1, iy);
                b = 0.25 * (fy(ix, iy + 1) + fy(ix, iy) + fy(ix + 1, iy + 1) + fy(ix + 1, iy));
            } else if (xh) {
                a = fx(ix, iy);
                b = 0.25 * (fy(ix, iy + 1) + fy(ix, iy) + fy(ix - 1, iy + 1) + fy(ix - 1, iy));
            } else if (yl) {
                a = 0.25 * (fx(ix + 1, iy) + fx(ix, iy) + fx(ix + 1, iy + 1) + fx(ix, iy + 1));
                b = fy(ix, iy + 1);
            } else if (yh) {
                a = 0.25 * (fx(ix + 1, iy) + fx(ix, iy) + fx(ix + 1, iy - 1) + fx(ix, iy - 1));
                b = fy(ix, iy);
            } else {
                a = 0.5 * (fx(ix + 1, iy) + fx(ix, iy));
                b = 0.5 * (fy(ix, iy + 1) + fy(ix, iy));
            }
            ux[(size_t)(iy - 1) * L.nox + ix - 1] = a;
            uy[(size_t)(iy - 1) * L.nox + ix - 1] = b;
        }
}

// strtod over exactly n numbers, then only blanks (Restart.cpp:182-193)
bool parse_row(const std::string &line, double *v, int n) {
    const char *p = line.c_str();
    char *e;
    for (int k = 0; k < n; ++k) {
        v[k] = std::strtod(p, &e);
        if (e == p || !std::isfinite(v[k])) return false;
        p = e;
    }
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    return *p == '\0';
}

}  // namespace

extern "C" {

int yz_write_dat(const char *path, const yz_case *c, const double *h, const double *s, const double *cc,
                 const double *ux, const double *uy, double L0, double H0, double U0, int dimensioned) {
    Layout L;
    if (!path || !layout_of(c, L) || !h || !s || !cc || !ux || (L.dim == 2 && !uy)) {
        yz::set_error("yz_write_dat: bad argument");
        return 2;
    }
    std::vector<double> hh = points_of(L, h), ss = points_of(L, s), co = points_of(L, cc), vx, vy;
    velocity_points(L, ux, uy, vx, vy);
    std::vector<double> xx = coords(L.nox, L.xmin, L.xmax, L.dx);
    std::vector<double> yy = L.dim == 2 ? coords(L.noy, L.ymin, L.ymax, L.dy) : std::vector<double>();
    if (dimensioned) {
        for (double &x : xx) x *= L0 * 1.0e3;
        for (double &y : yy) y *= L0 * 1.0e3;
        for (double &v : hh) v *= H0 * 1e6;
        for (double &v : ss) v *= H0 * 1e6;
        for (double &v : vx) v *= U0;
        for (double &v : vy) v *= U0;
    }
    const std::string fpath = path, tmp = fpath + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) {
        yz::set_error("yz_write_dat: cannot open the file");
        return 2;
    }
    if (L.dim == 1) {
        std::fputs(dimensioned ? "xx[mm] \t HS[um] \t Ux[m/s] \t CC[-] \t HH[um] \t SS[um]\n"
                               : "xx[-] \t HS[-] \t Ux[-] \t CC[-] \t HH[-] \t SS[-]\n", f);
        for (int i = 0; i < L.nox; ++i)
            std::fprintf(f, "%.17e\t%.17e\t%.17e\t%.17e\t%.17e\t%.17e\n", xx[i], hh[i] + ss[i], vx[i], co[i], hh[i],
                         ss[i]);
    } else {
        std::fputs(dimensioned ? "xx[mm] \t yy[mm] \t HS[um] \t Ux[m/s] \t Uy[m/s] \t CC[-] \t HH[um] \t SS[um]\n"
                               : "xx[-] \t yy[-] \t HS[-] \t Ux[-] \t Uy[-] \t CC[-] \t HH[-] \t SS[-]\n", f);
        for (int j = 0; j < L.noy; ++j)
            for (int i = 0; i < L.nox; ++i) {
                const size_t k = (size_t)j * L.nox + i;
                std::fprintf(f, "%.17e\t%.17e\t%.17e\t%.17e\t%.17e\t%.17e\t%.17e\t%.17e\n", xx[i], yy[j],
                             hh[k] + ss[k], vx[k], vy[k], co[k], hh[k], ss[k]);
            }
    }
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad || std::rename(tmp.c_str(), fpath.c_str()) != 0) {
        std::remove(tmp.c_str());
        yz::set_error("yz_write_dat: write failed");
        return 2;
    }
    return 0;
}

int yz_read_dat(const char *path, const yz_case *c, double H_coef, double *h, double *s, double *cc) {
    Layout L;
    if (!path || !layout_of(c, L) || !h || !s || !cc) {
        yz::set_error("yz_read_dat: bad argument");
        return 2;
    }
    std::ifstream fin(path);
    if (!fin.good()) {
        yz::set_error("yz_read_dat: cannot open the file");
        return 2;
    }
    const int ncol = L.dim == 2 ? 8 : 6;
    std::string line;
    if (!std::getline(fin, line)) {
        yz::set_error("yz_read_dat: empty file");
        return 2;
    }
    // a file with the driving potential carries one more column
    const int nread = line.find("PHI[") != std::string::npos ? ncol + 1 : ncol;
    std::vector<std::vector<double>> rows;
    while (std::getline(fin, line)) {
        std::vector<double> v(nread);
        if (!parse_row(line, v.data(), nread)) {
            yz::set_error("yz_read_dat: a row is not the expected count of finite numbers");
            return 2;
        }
        rows.push_back(v);
    }
    if ((long)rows.size() != (long)L.nox * L.noy) {
        yz::set_error("yz_read_dat: the file's point count does not match the grid");
        return 2;
    }
    for (size_t k = 0; k < rows.size(); ++k) {
        const std::vector<double> &v = rows[k];
        int p;
        if (L.dim == 1) {
            p = (int)k + 1;
            h[p] = v[4] * H_coef; s[p] = v[5] * H_coef; cc[p] = v[3];
        } else {
            const int ix = (int)(k % L.nox) + 1, iy = (int)(k / L.nox) + 1;
            p = iy * L.nxp + ix;
            h[p] = v[6] * H_coef; s[p] = v[7] * H_coef; cc[p] = v[5];
        }
    }
    return 0;
}

// The VTK structured-grid file (.vts) of a well-mixed film
// (OutputVTK::Write, Output.cpp:375-637, of Field_WM1D_Out / Field_WM2D_Out
// as ConvertFromStg fills them, LubFields.cpp:370-421, 832-880): the film
// drawn as two layers of the same points, the substrate surface (k = 0) and
// the free surface (k = 1). An XML header names nine point arrays and the
// coordinates, each an offset into the appended raw data that follows "_":
// every block is a 4-byte length, then little-endian doubles, x fastest,
// then y, then the layer. Scalars repeat the layer-0 value on layer 1;
// velocities are (ux, uy, 0) at both layers; a point's z is s + (h+s - s) k.
int yz_write_vts(const char *path, const yz_case *c, double time, const double *h, const double *s, const double *cc,
                 const double *ux6, const double *uy6, double L0, double H0, double U0, double T0, int dimensioned) {
    Layout L;
    if (!path || !layout_of(c, L) || !h || !s || !cc || !ux6 || (L.dim == 2 && !uy6)) {
        yz::set_error("yz_write_vts: bad argument");
        return 2;
    }
    const size_t np = (size_t)L.nxp * L.nyp, m = (size_t)L.nox * L.noy;
    std::vector<double> hh = points_of(L, h), ss = points_of(L, s), co = points_of(L, cc), hs(m);
    for (size_t k = 0; k < m; ++k) hs[k] = hh[k] + ss[k];
    // the mixed velocity, then the capillary, Marangoni, vertical-gravity,
    // in-plane gravity and centrifugal terms' own fields
    std::vector<double> vx[6], vy[6];
    for (int f = 0; f < 6; ++f)
        velocity_points(L, ux6 + f * np, L.dim == 2 ? uy6 + f * np : nullptr, vx[f], vy[f]);
    std::vector<double> xx = coords(L.nox, L.xmin, L.xmax, L.dx);
    std::vector<double> yy = L.dim == 2 ? coords(L.noy, L.ymin, L.ymax, L.dy) : std::vector<double>(1, 0.0);
    if (dimensioned) {
        time *= T0;
        for (double &x : xx) x *= L0 * 1.0e3;
        if (L.dim == 2)
            for (double &y : yy) y *= L0 * 1.0e3;
        for (size_t k = 0; k < m; ++k) {
            ss[k] *= H0 * 1e6;
            hs[k] *= H0 * 1e6;
        }
        for (int f = 0; f < 6; ++f)
            for (size_t k = 0; k < m; ++k) {
                vx[f][k] *= U0;
                vy[f][k] *= U0;
            }
    }
    const unsigned nz = 2, size_s = (unsigned)(m * nz * 8), size_v = 3 * size_s, step_s = size_s + 4,
                   step_v = size_v + 4;
    char line[512];
    std::string out;
    auto put = [&](const char *fmt, auto... a) {
        std::snprintf(line, sizeof line, fmt, a...);
        out += line;
    };
    out += "<VTKFile type=\"StructuredGrid\" GhostLevel=\"0\" byte_order=\"LittleEndian\">\n";
    put("<YuzenData Time=\"%e\" Model=\"WM2D\" Nx=\"%d\" Ny=\"%d\" Nz=\"%d\" />\n", time, L.nox, L.noy, (int)nz);
    put("<StructuredGrid WholeExtent=\"1 %d 1 %d 1 %d\">\n", L.nox, L.noy, (int)nz);
    put("<Piece Extent=\"1 %d 1 %d 1 %d\">\n", L.nox, L.noy, (int)nz);
    out += "<PointData>\n";
    unsigned off = 0;
    for (const char *name : {"Conc", "Height", "Substrate"}) {
        put("<DataArray Name=\"%s\" type=\"Float64\" format=\"appended\" offset=\"%u\"/>\n", name, off);
        off += step_s;
    }
    for (const char *name : {"Velocity", "V_Ca", "V_Ma", "V_Bo", "V_Pg", "V_Sp"}) {
        put("<DataArray Name=\"%s\" type=\"Float64\"  NumberOfComponents=\"3\" format=\"appended\" offset=\"%u\"/>\n",
            name, off);
        off += step_v;
    }
    out += "</PointData>\n<Points>\n";
    put("<DataArray Name=\"Coordinate\" type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"%u\"/>\n",
        off);
    out += "</Points>\n</Piece>\n</StructuredGrid>\n<AppendedData encoding=\"raw\">\n   _";
    auto raw = [&](const void *p, size_t bytes) { out.append(static_cast<const char *>(p), bytes); };
    for (const std::vector<double> *v : {&co, &hs, &ss}) {
        raw(&size_s, 4);
        for (unsigned k = 0; k < nz; ++k) raw(v->data(), m * 8);
    }
    for (int f = 0; f < 6; ++f) {
        raw(&size_v, 4);
        for (unsigned kz = 0; kz < nz; ++kz)
            for (size_t k = 0; k < m; ++k) {
                const double u[3] = {vx[f][k], L.dim == 2 ? vy[f][k] : 0.0, 0.0};
                raw(u, sizeof u);
            }
    }
    raw(&size_v, 4);
    for (unsigned kz = 0; kz < nz; ++kz) {
        const double zeta = (double)kz;
        for (int j = 0; j < L.noy; ++j)
            for (int i = 0; i < L.nox; ++i) {
                const size_t k = (size_t)j * L.nox + i;
                const double p3[3] = {xx[i], yy[j], ss[k] + (hs[k] - ss[k]) * zeta};
                raw(p3, sizeof p3);
            }
    }
    out += "\n</AppendedData>\n</VTKFile>\n";
    const std::string fpath = path, tmp = fpath + ".tmp";
    FILE *fo = std::fopen(tmp.c_str(), "wb");
    if (!fo) {
        yz::set_error("yz_write_vts: cannot open the file");
        return 2;
    }
    const bool bad = std::fwrite(out.data(), 1, out.size(), fo) != out.size();
    if (std::fclose(fo) != 0 || bad || std::rename(tmp.c_str(), fpath.c_str()) != 0) {
        std::remove(tmp.c_str());
        yz::set_error("yz_write_vts: write failed");
        return 2;
    }
    return 0;
}

}  // extern "C"
