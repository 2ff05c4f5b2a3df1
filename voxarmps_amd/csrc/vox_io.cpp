// vox_io.cpp — .grd particle input (reader of main.cu:496-682), ASCII VTU
// output (writers of inOut.cpp:104-336) and VTU restart input (getTP /
// getVTU, inOut.cpp:25-102) for libvoxarmps. Host code only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vox_internal.hpp"

using namespace vox;

extern "C" {

int vox_grd_count(const char* path, int* tp)
{
    if (!path || !tp) return fail(VOX_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(VOX_ERR_ARG, std::string("cannot open ") + path);
    int n = 0;
    int ok = std::fscanf(f, "%d", &n);
    std::fclose(f);
    if (ok != 1 || n <= 0) return fail(VOX_ERR_ARG, "bad particle count line");
    *tp = n;
    return VOX_OK;
}

// Rows: 2D "x y PTYPE u v p", 3D "x y z PTYPE u v w p" (main.cu:640-668).
int vox_grd_read(const char* path, int dim, int tp, double* x, double* y, double* z, double* u,
                 double* v, double* w, double* p, int32_t* ptype, int* gp, int* wp, int* fp)
{
    if (!path || !x || !y || !u || !v || !p || !ptype || (dim == 3 && (!z || !w)))
        return fail(VOX_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(VOX_ERR_ARG, std::string("cannot open ") + path);
    int n = 0;
    if (std::fscanf(f, "%d", &n) != 1 || n != tp) {
        std::fclose(f);
        return fail(VOX_ERR_ARG, "particle count differs from tp");
    }
    int g = 0, wl = 0, fl = 0;
    for (int i = 0; i < tp; ++i) {
        int t = 0, ok;
        if (dim == 3)
            ok = std::fscanf(f, "%lf %lf %lf %d %lf %lf %lf %lf", &x[i], &y[i], &z[i], &t, &u[i],
                             &v[i], &w[i], &p[i]) == 8;
        else
            ok = std::fscanf(f, "%lf %lf %d %lf %lf %lf", &x[i], &y[i], &t, &u[i], &v[i], &p[i]) == 6;
        if (!ok) {
            std::fclose(f);
            return fail(VOX_ERR_ARG, "malformed row " + std::to_string(i + 1));
        }
        ptype[i] = t;
        if (t < 0) ++g;                                       // main.cu:659-661
        else if (t == 0) ++wl;
        else ++fl;
    }
    std::fclose(f);
    if (gp) *gp = g;
    if (wp) *wp = wl;
    if (fp) *fp = fl;
    return VOX_OK;
}

int vox_grd_write(const char* path, int dim, int tp, const double* x, const double* y,
                  const double* z, const double* u, const double* v, const double* w,
                  const double* p, const int32_t* ptype)
{
    if (!path || !x || !y || !u || !v || !p || !ptype || (dim == 3 && (!z || !w)))
        return fail(VOX_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(VOX_ERR_ARG, std::string("cannot write ") + path);
    std::fprintf(f, "%d\n", tp);
    for (int i = 0; i < tp; ++i) {
        if (dim == 3)
            std::fprintf(f, "%.17g %.17g %.17g %d %.17g %.17g %.17g %.17g\n", x[i], y[i], z[i],
                         ptype[i], u[i], v[i], w[i], p[i]);
        else
            std::fprintf(f, "%.17g %.17g %d %.17g %.17g %.17g\n", x[i], y[i], ptype[i], u[i], v[i],
                         p[i]);
    }
    std::fclose(f);
    return VOX_OK;
}

// UnstructuredGrid of vertex cells; Position/Velocity with 3 components
// (z and w written in 3D), Pressure, partType (inOut.cpp:118-215). Values are
// written as Float64 with 17 significant digits, so a file read back by
// vox_vtu_read restores the state bit for bit; the line layout is the one the
// reference's getVTU (inOut.cpp:43-102) skips through.
int vox_write_vtu(vox_ctx* c, const char* path, int fluid_only)
{
    if (!c || !path) return fail(VOX_ERR_ARG, "null argument");
    const int n = c->tp, dim = c->p.dim;
    std::vector<double> x(n), y(n), z(n, 0.0), u(n), v(n), w(n, 0.0), p(n);
    std::vector<int32_t> t(n);
    VOX_TRY(vox_download(c, VOX_F_X, x.data()));
    VOX_TRY(vox_download(c, VOX_F_Y, y.data()));
    VOX_TRY(vox_download(c, VOX_F_U, u.data()));
    VOX_TRY(vox_download(c, VOX_F_V, v.data()));
    VOX_TRY(vox_download(c, VOX_F_P, p.data()));
    VOX_TRY(vox_download(c, VOX_F_PTYPE, t.data()));
    if (dim == 3) {
        VOX_TRY(vox_download(c, VOX_F_Z, z.data()));
        VOX_TRY(vox_download(c, VOX_F_W, w.data()));
    }
    std::vector<int> sel;
    for (int i = 0; i < n; ++i)
        if (!fluid_only || t[i] > 0) sel.push_back(i);
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(VOX_ERR_ARG, std::string("cannot write ") + path);
    const int m = (int)sel.size();
    std::fprintf(f, "<?xml version='1.0' encoding='UTF-8'?>\n");
    std::fprintf(f, "<VTKFile xmlns='VTK' byte_order='LittleEndian' version='0.1' type='UnstructuredGrid'>\n");
    std::fprintf(f, " <UnstructuredGrid>\n  <Piece NumberOfCells='%d' NumberOfPoints='%d'>\n", m, m);
    std::fprintf(f, "   <Points>\n    <DataArray NumberOfComponents='3' type='Float64' Name='Position' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.17g %.17g %.17g ", x[i], y[i], z[i]);
    std::fprintf(f, "\n    </DataArray>\n   </Points>\n   <PointData>\n");
    std::fprintf(f, "    <DataArray NumberOfComponents='3' type='Float64' Name='Velocity' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.17g %.17g %.17g ", u[i], v[i], w[i]);
    std::fprintf(f, "\n    </DataArray>\n");
    std::fprintf(f, "    <DataArray NumberOfComponents='1' type='Float64' Name='Pressure' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.17g ", p[i]);
    std::fprintf(f, "\n    </DataArray>\n");
    std::fprintf(f, "    <DataArray NumberOfComponents='1' type='Int32' Name='partType' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%d ", t[i]);
    std::fprintf(f, "\n    </DataArray>\n   </PointData>\n   <Cells>\n");
    std::fprintf(f, "    <DataArray type='Int32' Name='connectivity' format='ascii'>\n");
    for (int k = 0; k < m; ++k) std::fprintf(f, "%d ", k);
    std::fprintf(f, "\n    </DataArray>\n    <DataArray type='Int32' Name='offsets' format='ascii'>\n");
    for (int k = 1; k <= m; ++k) std::fprintf(f, "%d ", k);
    std::fprintf(f, "\n    </DataArray>\n    <DataArray type='UInt8' Name='types' format='ascii'>\n");
    for (int k = 0; k < m; ++k) std::fprintf(f, "1 ");
    std::fprintf(f, "\n    </DataArray>\n   </Cells>\n  </Piece>\n </UnstructuredGrid>\n</VTKFile>");
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return fail(VOX_ERR_ARG, std::string("write failed: ") + path);
    return VOX_OK;
}

static int read_text(const char* path, std::string& out)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(VOX_ERR_ARG, std::string("cannot open ") + path);
    char buf[1 << 16];
    size_t k;
    out.clear();
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    std::fclose(f);
    return VOX_OK;
}

// Number of points of the Piece (getTP, inOut.cpp:25-41, reads the first
// quoted number of the Piece line).
static int vtu_points(const std::string& s, int* tp)
{
    const size_t a = s.find("<Piece");
    if (a == std::string::npos) return fail(VOX_ERR_ARG, "VTU without a Piece element");
    const size_t b = s.find("NumberOfPoints='", a);
    if (b == std::string::npos) return fail(VOX_ERR_ARG, "VTU Piece without NumberOfPoints");
    const long n = std::strtol(s.c_str() + b + 16, nullptr, 10);
    if (n <= 0 || n > (1L << 31) - 1) return fail(VOX_ERR_ARG, "bad NumberOfPoints");
    *tp = (int)n;
    return VOX_OK;
}

// The body of DataArray `name`: count numbers parsed with strtod.
static int vtu_array(const std::string& s, const char* name, long count, std::vector<double>& out)
{
    const std::string key = std::string("Name='") + name + "'";
    const size_t a = s.find(key);
    if (a == std::string::npos) return fail(VOX_ERR_ARG, std::string("VTU without DataArray ") + name);
    const size_t b = s.find('>', a);
    if (b == std::string::npos) return fail(VOX_ERR_ARG, "truncated VTU");
    out.resize(count);
    const char* q = s.c_str() + b + 1;
    for (long k = 0; k < count; ++k) {
        char* e = nullptr;
        const double v = std::strtod(q, &e);
        if (e == q) return fail(VOX_ERR_ARG, std::string("DataArray ") + name + " has fewer than " +
                                                 std::to_string(count) + " values");
        out[k] = v;
        q = e;
    }
    return VOX_OK;
}

int vox_vtu_count(const char* path, int* tp)
{
    if (!path || !tp) return fail(VOX_ERR_ARG, "null argument");
    std::string s;
    VOX_TRY(read_text(path, s));
    return vtu_points(s, tp);
}

// Restart from a VTU (getVTU, inOut.cpp:43-102): positions, velocities,
// pressure and type of tp points, and the type counts (ghost, wall, fluid).
int vox_vtu_read(const char* path, int dim, int tp, double* x, double* y, double* z, double* u,
                 double* v, double* w, double* p, int32_t* ptype, int* gp, int* wp, int* fp)
{
    if (!path || !x || !y || !u || !v || !p || !ptype || (dim == 3 && (!z || !w)))
        return fail(VOX_ERR_ARG, "null argument");
    std::string s;
    VOX_TRY(read_text(path, s));
    int n = 0;
    VOX_TRY(vtu_points(s, &n));
    if (n != tp) return fail(VOX_ERR_ARG, "VTU point count differs from tp");
    std::vector<double> pos, vel, pr, ty;
    VOX_TRY(vtu_array(s, "Position", 3L * tp, pos));
    VOX_TRY(vtu_array(s, "Velocity", 3L * tp, vel));
    VOX_TRY(vtu_array(s, "Pressure", tp, pr));
    VOX_TRY(vtu_array(s, "partType", tp, ty));
    int g = 0, wl = 0, fl = 0;
    for (int i = 0; i < tp; ++i) {
        x[i] = pos[3L * i]; y[i] = pos[3L * i + 1];
        u[i] = vel[3L * i]; v[i] = vel[3L * i + 1];
        if (dim == 3) { z[i] = pos[3L * i + 2]; w[i] = vel[3L * i + 2]; }
        p[i] = pr[i];
        const int t = (int)ty[i];
        ptype[i] = t;
        if (t < 0) ++g;                                       // inOut.cpp:95-97
        else if (t == 0) ++wl;
        else ++fl;
    }
    if (gp) *gp = g;
    if (wp) *wp = wl;
    if (fp) *fp = fl;
    return VOX_OK;
}

}  // extern "C"
