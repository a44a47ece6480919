// Host-side mesh construction for every multigrid level: the arrays the
// reference builds once in read_instructions (Instructions.c:94-136) and that
// the hot path only reads.  The arithmetic (float storage, double
// accumulation, evaluation order) follows the cited reference routines so the
// uploaded geometry is bit-identical to the reference's E-> arrays.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "chip_common.h"
#include "../../include/citcom_learn.h"
#include "chip_kernels.h"

namespace chip {

namespace {

constexpr double BETA = 0.57735026918962576451;
// local node a (1..8) -> offsets (z, x, y) of element_definitions.h:214-222
const int offz[9] = {0, 0, 0, 0, 0, 1, 1, 1, 1};
const int offx[9] = {0, 0, 1, 1, 0, 0, 1, 1, 0};
const int offy[9] = {0, 0, 0, 1, 1, 0, 0, 1, 1};
const int bb[3][9] = {{0, 1, 2, 2, 1, 1, 2, 2, 1}, {0, 1, 1, 2, 2, 1, 1, 2, 2}, {0, 1, 1, 1, 1, 2, 2, 2, 2}};
const int eps[4][4] = {{0, 0, 0, 0}, {0, 1, -1, 1}, {0, -1, 1, -1}, {0, 1, -1, 1}};

inline int gnv(int n, int v) { return 8 * (n - 1) + (v - 1); }
inline int gnvx(int d, int n, int v) { return 64 * d + 8 * (n - 1) + (v - 1); }
inline int gnpx(int d, int n) { return 8 * d + (n - 1); }

double lpoly(int p, double y) { return p == 1 ? 0.5 * (1 - y) : (p == 2 ? 0.5 * (1 + y) : 0.0); }
double lpolydash(int p) { return p == 1 ? -0.5 : (p == 2 ? 0.5 : 0.0); }

struct Master {
    float gp[9][3];
    double Nv[64];
    float Nxv[192], Nxp[24];
};

// construct_shape_functions, Shape_functions.c:53-170 (3-D)
void master(Master &m) {
    const float b = (float)BETA;
    const float s[9][3] = {{0, 0, 0}, {-b, -b, -b}, {b, -b, -b}, {b, b, -b}, {-b, b, -b},
                           {-b, -b, b}, {b, -b, b}, {b, b, b}, {-b, b, b}};
    std::memcpy(m.gp, s, sizeof s);
    for (int i = 1; i <= 8; i++) {
        for (int j = 1; j <= 8; j++) {
            m.Nv[gnv(i, j)] = 1.0;
            for (int d = 1; d <= 3; d++) m.Nv[gnv(i, j)] *= lpoly(bb[d - 1][i], m.gp[j][d - 1]);
            for (int dd = 1; dd <= 3; dd++) {
                float v = (float)lpolydash(bb[dd - 1][i]);
                for (int d = 1; d <= 3; d++)
                    if (d != dd) v = (float)((double)v * lpoly(bb[d - 1][i], m.gp[j][d - 1]));
                m.Nxv[gnvx(dd - 1, i, j)] = v;
            }
        }
        for (int dd = 1; dd <= 3; dd++) {
            float v = (float)lpolydash(bb[dd - 1][i]);
            for (int d = 1; d <= 3; d++)
                if (d != dd) v = (float)((double)v * lpoly(bb[d - 1][i], 0.0));
            m.Nxp[gnpx(dd - 1, i)] = v;
        }
    }
}

double det3(double A[4][4]) {
    return (A[1][1] * (A[2][2] * A[3][3] - A[2][3] * A[3][2]) - A[1][2] * (A[2][1] * A[3][3] - A[2][3] * A[3][1]) +
            A[1][3] * (A[2][1] * A[3][2] - A[2][2] * A[3][1]));
}

// cofactor, General_matrix_functions.c:1373-1404
double cof3(double A[4][4], int i, int j) {
    double B[4][4];
    int p = 1, q = 1;
    for (int k = 1; k <= 3; k++) {
        if (k == i) continue;
        for (int l = 1; l <= 3; l++) {
            if (l == j) continue;
            B[p][q] = A[k][l];
            q++;
        }
        q = 1;
        p++;
    }
    return eps[i][j] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]);
}

struct HostLevel {
    Geo g;
    std::vector<float> X[3];
    std::vector<unsigned> NODE;
    std::vector<float> GNX, GDA, ECO, TWW, MASS, iw;
};

inline int nid(const Geo &g, int k, int i, int j) { return k + g.noz * (i + g.nox * j); }

// get_global_shape_fn (Size_does_matter.c:60-179) for every element, then
// mass_matrix (Size_does_matter.c:618-757), cartesian branches
void geometry(const Master &m, HostLevel &h) {
    const Geo &g = h.g;
    h.GNX.assign((size_t)g.nel * GNX_SZ, 0.0f);
    h.GDA.assign((size_t)g.nel * GDA_SZ, 0.0f);
    h.ECO.assign((size_t)g.nel * ECO_SZ, 0.0f);
    h.TWW.assign((size_t)g.nel * TWW_SZ, 0.0f);
    h.MASS.assign(g.nno, 0.0f);
    for (int e = 0; e < g.nel; e++) {
        const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
        int nodes[9];
        for (int a = 1; a <= 8; a++) nodes[a] = nid(g, ez + offz[a], ex + offx[a], ey + offy[a]);
        float *gnx = &h.GNX[(size_t)e * GNX_SZ];
        float *gda = &h.GDA[(size_t)e * GDA_SZ];
        for (int k = 1; k <= 9; k++) {  // k = 9: the pressure point
            const int kk = (k == 9) ? 1 : k;
            double dxda[4][4], cof[4][4];
            for (int d = 1; d <= 3; d++)
                for (int ee = 1; ee <= 3; ee++) dxda[d][ee] = 0.0;
            for (int i = 1; i <= 8; i++)
                for (int d = 1; d <= 3; d++)
                    for (int ee = 1; ee <= 3; ee++) {
                        const float nx = (k == 9) ? m.Nxp[gnpx(d - 1, i)] : m.Nxv[gnvx(d - 1, i, kk)];
                        const float pr = h.X[ee - 1][nodes[i]] * nx;
                        dxda[d][ee] += pr;
                    }
            const double jac = det3(dxda);
            if (k == 9) gda[10] = (float)jac;
            else gda[k] = (float)jac;
            for (int d = 1; d <= 3; d++)
                for (int ee = 1; ee <= 3; ee++) cof[d][ee] = cof3(dxda, d, ee);
            for (int j = 1; j <= 8; j++)
                for (int d = 1; d <= 3; d++) {
                    float v = 0.0f;
                    for (int ee = 1; ee <= 3; ee++) {
                        const float nx = (k == 9) ? m.Nxp[gnpx(ee - 1, j)] : m.Nxv[gnvx(ee - 1, j, kk)];
                        v = (float)((double)v + nx * cof[ee][d]);
                    }
                    v = (float)((double)v / jac);
                    if (k == 9) gnx[192 + gnpx(d - 1, j)] = v;
                    else gnx[gnvx(d - 1, j, kk)] = v;
                }
        }
        // mass_matrix
        float *eco = &h.ECO[(size_t)e * ECO_SZ];
        double centre[4] = {0, 0, 0, 0}, area = 0.0, temp[9];
        for (int i = 1; i <= 3; i++) {
            for (int a = 1; a <= 8; a++) centre[i] += h.X[i - 1][nodes[a]];
            eco[i] = (float)(centre[i] / 8);
        }
        const std::vector<float> &X0 = h.X[0], &X1 = h.X[1], &X2 = h.X[2];
        const float dx2 = (float)(0.25 * (X1[nodes[3]] + X1[nodes[4]] + X1[nodes[7]] + X1[nodes[8]] - X1[nodes[1]] -
                                          X1[nodes[2]] - X1[nodes[5]] - X1[nodes[6]]));
        eco[6] = (float)std::sqrt(dx2 * dx2);
        const float dx1 = (float)(0.25 * (X0[nodes[2]] + X0[nodes[3]] + X0[nodes[6]] + X0[nodes[7]] - X0[nodes[1]] -
                                          X0[nodes[4]] - X0[nodes[5]] - X0[nodes[8]]));
        eco[5] = (float)std::sqrt(dx1 * dx1);
        const float dx3 = (float)(0.25 * (X2[nodes[5]] + X2[nodes[6]] + X2[nodes[7]] + X2[nodes[8]] - X2[nodes[1]] -
                                          X2[nodes[2]] - X2[nodes[3]] - X2[nodes[4]]));
        eco[7] = (float)std::sqrt(dx3 * dx3);
        for (int n = 1; n <= 8; n++) {
            const float w = 1.0f * gda[n];
            area += w;
        }
        eco[12] = (float)area;
        for (int a = 1; a <= 8; a++) {
            temp[a] = 0.0;
            for (int n = 1; n <= 8; n++) {
                const float w = gda[n] * 1.0f;
                temp[a] += w * m.Nv[gnv(a, n)];
            }
        }
        for (int a = 1; a <= 8; a++) h.MASS[nodes[a]] = (float)((double)h.MASS[nodes[a]] + temp[a]);
        for (int a = 1; a <= 8; a++) h.TWW[(size_t)e * TWW_SZ + a] = (float)temp[a];
    }
    for (int n = 0; n < g.nno; n++) h.MASS[n] = (float)(1.0 / h.MASS[n]);
}

// velocity boundary flags of every level (Boundary_conditions.c:44-156 with
// velocity_refl_vert_bc :249-372 and horizontal_bc :602-643)
void velocity_flags(const chip_problem &p, HostLevel &h) {
    const Geo &g = h.g;
    h.NODE.assign(g.nno, 0u);
    for (int j = 0; j < g.noy; j++)
        for (int i = 0; i < g.nox; i++)
            for (int k = 0; k < g.noz; k++) {
                unsigned f = 0;
                const bool zb = (k == 0 || k == g.noz - 1);
                if (zb) f |= VBZ;
                if (i == 0 || i == g.nox - 1) {
                    f |= VBX;
                    if (!zb) f &= ~(VBY | VBZ);
                }
                if (j == 0 || j == g.noy - 1) {
                    f |= VBY;
                    if (!zb) f &= ~VBZ;
                    if (i != 0 && i != g.nox - 1 && !zb) f &= ~VBX;
                }
                if ((k == 0 && p.botvbc == 1) || (k == g.noz - 1 && p.topvbc == 1)) f |= VBX | VBY | VBZ;
                h.NODE[nid(g, k, i, j)] = f;
            }
}

// interp_vector weights (Solver_multigrid.c:196-294): n1 = x2/(x1+x2) with x1,
// x2 the sizes of the lowest-numbered elements holding the two coarse nodes
void interp_weights(HostLevel &h) {
    const Geo &g = h.g;
    h.iw.assign(2 * (g.nox + g.noz + g.noy), 0.0f);
    auto first_el = [&](int k, int i, int j) {
        return (k > 0 ? k - 1 : 0) + g.elz * ((i > 0 ? i - 1 : 0) + g.elx * (j > 0 ? j - 1 : 0));
    };
    for (int i = 1; i < g.nox - 1; i += 2) {
        const float x1 = h.ECO[(size_t)first_el(0, i - 1, 0) * ECO_SZ + 5];
        const float x2 = h.ECO[(size_t)first_el(0, i + 1, 0) * ECO_SZ + 5];
        h.iw[2 * i] = x2 / (x1 + x2);
        h.iw[2 * i + 1] = x1 / (x1 + x2);
    }
    float *wz = &h.iw[2 * g.nox];
    for (int k = 1; k < g.noz - 1; k += 2) {
        const float x1 = h.ECO[(size_t)first_el(k - 1, 0, 0) * ECO_SZ + 7];
        const float x2 = h.ECO[(size_t)first_el(k + 1, 0, 0) * ECO_SZ + 7];
        wz[2 * k] = x2 / (x1 + x2);
        wz[2 * k + 1] = x1 / (x1 + x2);
    }
    float *wy = &h.iw[2 * (g.nox + g.noz)];
    for (int j = 1; j < g.noy - 1; j += 2) {
        const float x1 = h.ECO[(size_t)first_el(0, 0, j - 1) * ECO_SZ + 6];
        const float x2 = h.ECO[(size_t)first_el(0, 0, j + 1) * ECO_SZ + 6];
        wy[2 * j] = x2 / (x1 + x2);
        wy[2 * j + 1] = x1 / (x1 + x2);
    }
}

template <typename T>
int dev_alloc(chip_ctx *c, T **ptr, size_t count) {
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    CHIP_CHECK(hipMalloc(&p, bytes));
    CHIP_CHECK(hipMemset(p, 0, bytes));
    c->allocs.push_back(p);
    c->dev_bytes += (long)bytes;
    *ptr = static_cast<T *>(p);
    return 0;
}

template <typename T>
int dev_upload(chip_ctx *c, T **ptr, const std::vector<T> &v) {
    if (dev_alloc(c, ptr, v.size())) return -1;
    CHIP_CHECK(hipMemcpy(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

int alloc_coarse_net(chip_ctx *c) {
    if (c->cnet_w) return 0;
    const size_t nno = (size_t)c->L[c->levmin].g.nno;
    if (dev_alloc(c, &c->cnet_w, (size_t)CNET_MAX_LAYERS * 16 * 27 * 16) ||
        dev_alloc(c, &c->cnet_b, (size_t)CNET_MAX_LAYERS * 16) || dev_alloc(c, &c->cnet_ys, 3) ||
        dev_alloc(c, &c->cnet_logeta, nno) || dev_alloc(c, &c->cnet_keep, 3 * (size_t)c->L[c->levmax].g.nno))
        return -1;
    return 0;
}

int alloc_warm_start(chip_ctx *c) {
    if (c->Uprev) return 0;
    const size_t nno = (size_t)c->L[c->levmax].g.nno, nel = (size_t)c->L[c->levmax].g.nel;
    return (dev_alloc(c, &c->Uprev, 3 * nno) || dev_alloc(c, &c->Pprev, nel) || dev_alloc(c, &c->Th0, nno) ||
            dev_alloc(c, &c->Uprev2, 3 * nno) || dev_alloc(c, &c->Pprev2, nel) || dev_alloc(c, &c->D2u, 3 * nno) ||
            dev_alloc(c, &c->D2p, nel) || dev_alloc(c, &c->wn_d2pn, nno) ||
            dev_alloc(c, &c->Th1, nno) || dev_alloc(c, &c->wn_dT, nno) || dev_alloc(c, &c->wn_pn, nno) ||
            dev_alloc(c, &c->wn_dpn, nno) || dev_alloc(c, &c->wn_logeta, nno) || dev_alloc(c, &c->wn_x, 16 * nno) ||
            dev_alloc(c, &c->Ue, 3 * nno) || dev_alloc(c, &c->Pe, nel))
               ? -1
               : 0;
}

int alloc_warm_net(chip_ctx *c) {
    if (c->wn_w) return 0;
    const Geo &g = c->L[c->levmax].g;
    long nw = 0, nb = 0;
    cl_unet_params(&nw, &nb);
    const long ws = cl_unet_workspace(1, g.nox, g.noy, g.noz);
    char *wsp = nullptr;
    if (dev_alloc(c, &c->wn_w, (size_t)nw) || dev_alloc(c, &c->wn_b, (size_t)nb) ||
        dev_alloc(c, &c->wn_out, 16 * (size_t)g.nno) || dev_alloc(c, &wsp, (size_t)ws))
        return -1;
    c->wn_ws = wsp;
    return 0;
}

int build_mesh(chip_ctx *c) {
    const chip_problem &p = c->p;
    Master m;
    master(m);
    c->levmin = 0;
    c->levmax = p.levels - 1;
    if (p.levels < 2 || p.levels > MAXL) {
        set_error("levels must be in [2, 12]");
        return -1;
    }
    // the viscosity laws visc_from_T implements here (Viscosity_structures.c:491-737)
    if (p.tdepv && p.rheol != 0 && p.rheol != 3 && p.rheol != 11) {
        set_error("rheol " + std::to_string(p.rheol) + " is not supported (0, 3 and 11 are)");
        return -1;
    }
    if (p.mixed < 0 || p.mixed > 2) {
        set_error("mixed must be 0, 1 or 2");
        return -1;
    }
    if (p.num_mat < 1 || p.num_mat > 4) {
        set_error("num_mat must be in [1, 4]");
        return -1;
    }
    for (int lev = c->levmin; lev <= c->levmax; lev++) {
        HostLevel h;
        Geo &g = h.g;
        const int f = 1 << lev, s = 1 << (c->levmax - lev);
        g.nox = p.mgunitx * f + 1;
        g.noy = p.mgunity * f + 1;
        g.noz = p.mgunitz * f + 1;
        g.elx = g.nox - 1;
        g.ely = g.noy - 1;
        g.elz = g.noz - 1;
        g.nno = g.nox * g.noy * g.noz;
        g.nel = g.elx * g.ely * g.elz;
        // colour classes by index parities (k&1) | (i&1)<<1 | (j&1)<<2
        int off = 0;
        for (int cc = 0; cc < 8; cc++) {
            const int pk = cc & 1, pi = (cc >> 1) & 1, pj = (cc >> 2) & 1;
            const int nz = (g.noz - pk + 1) / 2, nx = (g.nox - pi + 1) / 2, ny = (g.noy - pj + 1) / 2;
            g.cs[cc] = off;
            g.cn[cc] = nz * nx * ny;
            g.cz[cc] = nz;
            g.cx[cc] = nx;
            off += (g.cn[cc] + 63) / 64 * 64;
        }
        g.cs[8] = off;
        g.pitch = off;
        // coordinates: per-axis arrays of the finest level, injected downwards
        // (Nodal_mesh.c:170-180, 244-247)
        for (int d = 0; d < 3; d++) h.X[d].resize(g.nno);
        for (int j = 0; j < g.noy; j++)
            for (int i = 0; i < g.nox; i++)
                for (int k = 0; k < g.noz; k++) {
                    const int n = nid(g, k, i, j);
                    h.X[0][n] = c->xc[i * s];
                    h.X[1][n] = c->yc[j * s];
                    h.X[2][n] = c->zc[k * s];
                }
        velocity_flags(p, h);
        geometry(m, h);
        interp_weights(h);
        Level &L = c->L[lev];
        L.g = g;
        if (lev == c->levmax) {
            // temperature flags and values on the finest level
            // (Boundary_conditions.c:188-246, 516-565)
            std::vector<float> TB(g.nno, 0.0f);
            for (int j = 0; j < g.noy; j++)
                for (int i = 0; i < g.nox; i++)
                    for (int k = 0; k < g.noz; k++) {
                        const int n = nid(g, k, i, j);
                        if (k == 0 && p.bottbc >= 1) {
                            h.NODE[n] |= TBZ;
                            TB[n] = p.bottbcval;
                        }
                        if (k == g.noz - 1 && p.toptbc >= 1) {
                            h.NODE[n] |= TBZ;
                            TB[n] = p.toptbcval;
                        }
                        if (i == 0 || i == g.nox - 1) h.NODE[n] |= FBX;
                        if (j == 0 || j == g.noy - 1) h.NODE[n] |= FBY;
                    }
            if (dev_upload(c, &c->TB, TB)) return -1;
            // material groups (construct_mat_group, Construct_arrays.c:742-829;
            // layers, Viscosity_structures.c:1389-1404)
            std::vector<int> mat(g.nel);
            for (int e = 0; e < g.nel; e++) {
                const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
                float x3 = 0.0f;
                for (int a = 1; a <= 8; a++) x3 += h.X[2][nid(g, ez + offz[a], ex + offx[a], ey + offy[a])];
                x3 = x3 / 8;
                int l = 0;
                for (int i = 0; i < p.num_mat - 1; i++)
                    if (x3 >= p.z_layer[i]) {
                        l = i + 1;
                        break;
                    }
                mat[e] = l ? l : p.num_mat;
            }
            if (dev_upload(c, &c->mat, mat)) return -1;
            // thermexp_factor = thermdiff_factor = 1 (Instructions.c:1062-1084)
            std::vector<float> ones(g.noz, 1.0f);
            if (dev_upload(c, &c->expansivity, ones) || dev_upload(c, &c->diffusivity, ones)) return -1;
            if (dev_upload(c, &c->xc_d, c->xc) || dev_upload(c, &c->yc_d, c->yc)) return -1;
            // diffusive step limit of std_timestep (Advection_diffusion.c:761-775)
            float dts = 1.0e8f;
            for (int e = 0; e < g.nel; e++)
                for (int d = 1; d <= 3; d++) {
                    const float sz = h.ECO[(size_t)e * ECO_SZ + 4 + d];
                    const float ts = sz * sz;
                    if (dts > ts) dts = ts;
                }
            c->diff_timestep = (float)(0.5 * dts);
        }
        if (dev_upload(c, &L.NODE, h.NODE) || dev_upload(c, &L.GNX, h.GNX) || dev_upload(c, &L.GDA, h.GDA) ||
            dev_upload(c, &L.ECO, h.ECO) || dev_upload(c, &L.TWW, h.TWW) || dev_upload(c, &L.MASS, h.MASS) ||
            dev_upload(c, &L.iw, h.iw))
            return -1;
        if (dev_alloc(c, &L.K, (size_t)KROW * g.pitch) || dev_alloc(c, &L.BI, 3 * (size_t)g.nno) ||
            dev_alloc(c, &L.BPI, g.nel) || dev_alloc(c, &L.EG, (size_t)g.nel * EG_SZ) ||
            dev_alloc(c, &L.EVI, (size_t)g.nel * 8) || dev_alloc(c, &L.VN, g.nno) ||
            dev_alloc(c, &L.vel, 3 * (size_t)g.nno) || dev_alloc(c, &L.res, 3 * (size_t)g.nno) ||
            dev_alloc(c, &L.fl, 3 * (size_t)g.nno) || dev_alloc(c, &L.del, 3 * (size_t)g.nno) ||
            dev_alloc(c, &L.AU, 3 * (size_t)g.nno))
            return -1;
        if (p.mixed) {
            if (dev_alloc(c, &L.velf, 3 * (size_t)g.nno) || dev_alloc(c, &L.resf, 3 * (size_t)g.nno) ||
                dev_alloc(c, &L.BIf, 3 * (size_t)g.nno) ||
                dev_alloc(c, &L.flf, 3 * (size_t)g.nno) || dev_alloc(c, &L.delf, 3 * (size_t)g.nno) ||
                dev_alloc(c, &L.AUf, 3 * (size_t)g.nno))
                return -1;
            // the FP16 stiffness of the two finest levels, whose colour
            // passes and matvecs are bandwidth-bound
            if (p.mixed == 2 && lev >= c->levmax - 1 && lev >= c->levmin + 2 &&
                dev_alloc(c, &L.Kh, (size_t)KROW * g.pitch))
                return -1;
        }
        if (dev_alloc(c, &L.xbar, 1)) return -1;
        // isoviscous until a viscosity is set
        std::vector<float> one((size_t)g.nel * 8, 1.0f);
        CHIP_CHECK(hipMemcpy(L.EVI, one.data(), one.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    {
        Level &C = c->L[c->levmin];
        if (affine_ok(C) && dev_alloc(c, &C.Mc, 9 * (size_t)C.g.nno * C.g.nno)) return -1;
        // the level above: affine form when its two smoother calls in
        // multi_grid use the same number of sweeps
        if (C.Mc && c->levmin + 1 < c->levmax && p.down_heavy == p.up_heavy) {
            Level &A = c->L[c->levmin + 1];
            if (affine_ok(A)) {
                const size_t n3 = 3 * (size_t)A.g.nno;
                if (dev_alloc(c, &A.Ma, n3 * n3) || dev_alloc(c, &A.Mb, n3 * n3) || dev_alloc(c, &A.tmp, n3) ||
                    (p.mixed && (dev_alloc(c, &A.Maf, n3 * n3) || dev_alloc(c, &A.Mbf, n3 * n3))) ||
                    dev_alloc(c, &A.tmpf, n3))
                    return -1;
                A.affine_sweeps = p.down_heavy;
            }
        }
    }
    // the FP32 cycles' exact coarsest-level solve: the largest level of at
    // most DIRECT_MAX_DOF unknowns with at least two levels below it and
    // the finest above it (17x17x9 nodes at 128x128x64 elements)
    if (p.mixed)
        for (int lev = c->levmax - 1; lev >= c->levmin + 2; lev--)
            if (3L * c->L[lev].g.nno <= DIRECT_MAX_DOF) {
                const size_t n3 = 3 * (size_t)c->L[lev].g.nno;
                if (dev_alloc(c, &c->dA, n3 * n3) || dev_alloc(c, &c->dAinv, n3 * n3) ||
                    dev_alloc(c, &c->dAh, n3 * n3) || dev_alloc(c, &c->dAh_stage, n3 * n3) ||
                    dev_alloc(c, &c->dAs, n3) || dev_alloc(c, &c->dAs_stage, n3) || dev_alloc(c, &c->dinfo, 2))
                    return -1;
                c->dlev = lev;
                break;
            }
    // the FP32 cycles' box smoother: the levels between the exact (or the
    // affine) level and the finest, larger than one workgroup's level
    if (p.mixed)
        for (int lev = c->levmax - 1; lev > c->levmin + 1 && lev > c->dlev; lev--) {
            Level &L = c->L[lev];
            if (L.g.nno <= 1024 || !box_part(L.g.noz, L.g.nox, L.g.noy, L.bp)) continue;
            L.nbox = L.bp.n[0] * L.bp.n[1] * L.bp.n[2];
            L.Kbox_half = L.Kh != nullptr;
            const size_t n = box_tile_elems(L);
            if (L.Kbox_half ? dev_alloc(c, (uint16_t **)&L.Kbox, n) : dev_alloc(c, (float **)&L.Kbox, n)) return -1;
            if (!L.tmpf && dev_alloc(c, &L.tmpf, 3 * (size_t)L.g.nno)) return -1;
        }
    const Geo &g = c->L[c->levmax].g;
    const size_t nno = g.nno, nel = g.nel, neq = 3 * nno;
    if (dev_alloc(c, &c->U, neq) || dev_alloc(c, &c->P, nel) || dev_alloc(c, &c->F, neq) ||
        dev_alloc(c, &c->T, nno) || dev_alloc(c, &c->Tdot, nno) || dev_alloc(c, &c->V, 3 * nno) ||
        dev_alloc(c, &c->buoy, nno) || dev_alloc(c, &c->T1, nno) || dev_alloc(c, &c->Tdot1, nno) ||
        dev_alloc(c, &c->Eres, nel * 8) || dev_alloc(c, &c->H, g.noz) || dev_alloc(c, &c->hf, nno) ||
        dev_alloc(c, &c->r0, nel) || dev_alloc(c, &c->r1, nel) || dev_alloc(c, &c->r2, nel) ||
        dev_alloc(c, &c->z0, nel) || dev_alloc(c, &c->z1, nel) || dev_alloc(c, &c->s1, nel) ||
        dev_alloc(c, &c->s2, nel) || dev_alloc(c, &c->Ah, neq) || dev_alloc(c, &c->u1, neq) ||
        dev_alloc(c, &c->sr, neq) || dev_alloc(c, &c->sd, neq) || dev_alloc(c, &c->partial, 4 * 1024 + 64) ||
        dev_alloc(c, &c->counter, 4) || dev_alloc(c, &c->dots, 32) || dev_alloc(c, &c->minmax, 4) ||
        dev_alloc(c, &c->xcd_ids, 16))
        return -1;
    CHIP_CHECK(hipHostMalloc((void **)&c->h_dots, 32 * sizeof(double), hipHostMallocDefault));
    CHIP_CHECK(hipHostMalloc((void **)&c->h_minmax, 4 * sizeof(unsigned), hipHostMallocDefault));
    return 0;
}

}  // namespace chip
