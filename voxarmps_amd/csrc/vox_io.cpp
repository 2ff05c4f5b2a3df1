// vox_io.cpp — .grd particle input (reader of main.cu:496-682) and ASCII VTU
// output (writers of inOut.cpp:104-336) for libvoxarmps. Host code only.
#include <cstdio>
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
// (z and w written in 3D), Pressure, partType (inOut.cpp:118-215).
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
    std::fprintf(f, "   <Points>\n    <DataArray NumberOfComponents='3' type='Float32' Name='Position' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.9g %.9g %.9g ", x[i], y[i], z[i]);
    std::fprintf(f, "\n    </DataArray>\n   </Points>\n   <PointData>\n");
    std::fprintf(f, "    <DataArray NumberOfComponents='3' type='Float32' Name='Velocity' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.9g %.9g %.9g ", u[i], v[i], w[i]);
    std::fprintf(f, "\n    </DataArray>\n");
    std::fprintf(f, "    <DataArray NumberOfComponents='1' type='Float32' Name='Pressure' format='ascii'>\n");
    for (int i : sel) std::fprintf(f, "%.9g ", p[i]);
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
    std::fclose(f);
    return VOX_OK;
}

}  // extern "C"
