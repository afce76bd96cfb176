// fg_io.hip -- host readers of the drivers' on-disk formats (no device code).
//
// Parameter files (Preprocess.c of each driver, e.g. 01_SS_Transport_Equation/
// 2D/01_CommonFiles/Preprocess.c:44-61, 03_Euler_Equations/2D/01_CommonFiles/
// Preprocess.c:45-74, 01_SS_Transport_Equation/3D/01_CommonFiles/
// Preprocess.c:45-83): one field per line, "value<TAB>:label", read in the
// driver's fixed order with fscanf("%s\t:%[^\n]").  Here: the value of
// each of the first n fields (the first whitespace-delimited token of the
// line; kind 'p' = the Euler driver's Preconditioner line, read up to the
// first tab, :54), returned as NUL-separated strings for the caller's typed
// field table.
//
// Mesh files (<title>_<nnodes>_<nel>.dat): nnodes, then per node ndim
// coordinates and nint integers (2D scalar: x y Type, Preprocess.c:78;
// Euler: x y rhoType v1Type v2Type eType, 03 Preprocess.c:89; 3D: x y z
// Type), then nel, then per element nvert vertices and etype_cols integers
// (2D: v0 v1 v2 Type, Preprocess.c:98).  The whole file is read once and
// parsed with strtod/strtol (the numbers fscanf's %lf / %d would give).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "fg_internal.h"

namespace {

bool slurp(const char *path, std::string &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? fread(&buf[0], 1, (size_t)n, f) : 0;
    fclose(f);
    return got == buf.size();
}

struct Cursor {
    const char *p, *end;
    bool num_d(double &v)
    {
        char *q;
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p >= end) return false;
        v = strtod(p, &q);
        if (q == p) return false;
        p = q;
        return true;
    }
    bool num_i(int &v)
    {
        char *q;
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p >= end) return false;
        const long x = strtol(p, &q, 10);
        if (q == p) return false;
        v = (int)x;
        p = q;
        return true;
    }
};

}  // namespace

extern "C" int fg_param_file_read(const char *path, const char *kinds, char *values, int cap,
                                  int *count)
{
    FG_ARG(path && kinds && values && cap > 0 && count, "path/kinds/values/count");
    std::string buf;
    if (!slurp(path, buf)) {
        fg_set_error("cannot read %s", path);
        return FG_ERR_ARG;
    }
    const char *p = buf.c_str(), *end = p + buf.size();
    int n = 0, used = 0;
    for (const char *k = kinds; *k; k++) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p >= end) break;
        const char *s = p;
        if (*k == 'p')
            while (p < end && *p != '\t' && *p != '\n') p++;
        else
            while (p < end && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;
        const int len = (int)(p - s);
        if (used + len + 1 > cap) {
            fg_set_error("parameter buffer too small");
            return FG_ERR_ARG;
        }
        memcpy(values + used, s, (size_t)len);
        values[used + len] = '\0';
        used += len + 1;
        n++;
        while (p < end && *p != '\n') p++;   // the label
    }
    *count = n;
    if (n < (int)strlen(kinds)) {
        fg_set_error("%s: %d of %d fields", path, n, (int)strlen(kinds));
        return FG_ERR_ARG;
    }
    return FG_OK;
}

extern "C" int fg_mesh_file_read(const char *path, int ndim, int nint, int nvert, int etype_cols,
                                 int *nnodes, int *nel, double *xyz, int *node_ints,
                                 int *vertex, int *elem_ints)
{
    FG_ARG(path && nnodes && nel && ndim >= 1 && ndim <= 3 && nint >= 0 && nvert >= 1 &&
               etype_cols >= 0,
           "path/sizes/format");
    std::string buf;
    if (!slurp(path, buf)) {
        fg_set_error("cannot read %s", path);
        return FG_ERR_ARG;
    }
    Cursor c{buf.c_str(), buf.c_str() + buf.size()};
    int nn = 0, ne = 0;
    FG_ARG(c.num_i(nn) && nn >= 0, "node count");
    const bool fill = xyz != nullptr;
    FG_ARG(!fill || *nnodes == nn, "nnodes differs from the file's");
    for (long long i = 0; i < nn; i++) {
        for (int d = 0; d < ndim; d++) {
            double v;
            FG_ARG(c.num_d(v), "node coordinate");
            if (fill) xyz[i * ndim + d] = v;
        }
        for (int d = 0; d < nint; d++) {
            int v;
            FG_ARG(c.num_i(v), "node integer");
            if (fill && node_ints) node_ints[i * nint + d] = v;
        }
    }
    FG_ARG(c.num_i(ne) && ne >= 0, "element count");
    FG_ARG(!fill || *nel == ne, "nel differs from the file's");
    for (long long e = 0; e < ne; e++) {
        for (int d = 0; d < nvert; d++) {
            int v;
            FG_ARG(c.num_i(v), "element vertex");
            if (fill && vertex) vertex[e * nvert + d] = v;   // as written (GMSH: 1-based)
        }
        for (int d = 0; d < etype_cols; d++) {
            int v;
            FG_ARG(c.num_i(v), "element integer");
            if (fill && elem_ints) elem_ints[e * etype_cols + d] = v;
        }
    }
    *nnodes = nn;
    *nel = ne;
    return FG_OK;
}
