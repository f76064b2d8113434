// TEST INFRASTRUCTURE: a driver around the reference's corner-point geometry
// (OCP_COORD, src/CornerGrid.cpp), compiled from the reference's sources by
// oracle/Makefile 'ref' into _ref/libocpcorner.so.  It writes the golden
// vectors the corner-point reader (opencaeporo_amd/cornergrid.py) is checked
// against; nothing here is part of the product.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
// the results are private members of OCP_COORD (Grid, a friend, reads them)
#define private public
#include "CornerGrid.hpp"
#undef private

// COORD ((nx+1)(ny+1) pillars x 6) and ZCORN (8 nx ny nz) in deck order.
// Per cell: v, depth, dx, dy, dz; per connection (OCP_COORD::connect, in its
// order): begin, end, direction (1 x, 2 y, 3 z), Ad_dd_begin, Ad_dd_end.
// Returns 0, or -1 when more than maxConn connections were found (nConn is
// set either way).
extern "C" int ref_corner_grid(int nx, int ny, int nz, const double* coord, const double* zcorn, double* v,
                               double* depth, double* dx, double* dy, double* dz, int64_t maxConn, int64_t* nConn,
                               int32_t* begin, int32_t* end, int32_t* dir, double* adB, double* adE)
{
    OCP_COORD c;
    c.Allocate(nx, ny, nz);
    const size_t nb = (size_t)nx * ny * nz;
    std::vector<double> co(coord, coord + (size_t)6 * (nx + 1) * (ny + 1));
    std::vector<double> zc(zcorn, zcorn + 8 * nb);
    c.InputData(co, zc);
    c.SetupCornerPoints();
    for (size_t n = 0; n < nb; n++) {
        v[n] = c.v[n];
        depth[n] = c.depth[n];
        dx[n] = c.dx[n];
        dy[n] = c.dy[n];
        dz[n] = c.dz[n];
    }
    *nConn = (int64_t)c.numConn;
    if ((int64_t)c.numConn > maxConn) return -1;
    for (size_t q = 0; q < c.numConn; q++) {
        const GeneralConnect& g = c.connect[q];
        begin[q] = (int32_t)g.begin;
        end[q] = (int32_t)g.end;
        dir[q] = (int32_t)g.directionType;
        adB[q] = g.Ad_dd_begin;
        adE[q] = g.Ad_dd_end;
    }
    return 0;
}
