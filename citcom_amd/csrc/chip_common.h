// Internal data model of libcitcom_hip: one struct per multigrid level, all
// arrays resident in HBM.  Node/element numbering follows the reference
// (z fastest); the stiffness is stored as full 27-neighbour rows of 3x3
// FP32 blocks, coefficient-major in 8-colour order so that both the matvec
// and every colour pass of the Gauss-Seidel smoother stream contiguous,
// 256-byte aligned runs of each coefficient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/citcom_hip.h"

namespace chip {

constexpr int MAXL = 12;
constexpr int KROW = 243;   // 3 rows x 27 neighbour nodes x 3 columns
// largest level (unknowns) solved exactly in the FP32 cycles: its dense
// inverse is 3 nno x 3 nno FP32 (243 MB at 8192), applied a few times per
// multigrid call, factored in FP64 at every rebuild
constexpr long DIRECT_MAX_DOF = 8192;
constexpr int GNX_SZ = 216; // Shape_function_dx: vpt[3*8*8] + ppt[3*8]
constexpr int GDA_SZ = 11;  // Shape_function_dA: vpt[9] + ppt[2]
constexpr int ECO_SZ = 13;  // COORD: centre[4] size[4] recip_size[4] area
constexpr int TWW_SZ = 9;   // FNODE: node[9]
constexpr int EG_SZ = 24;   // EG: g[24][1]

// node flag bits (global_defs.h:65-89)
constexpr unsigned VBX = 0x2, VBZ = 0x4, VBY = 0x8;
constexpr unsigned TBX = 0x10, TBZ = 0x20, TBY = 0x40;
constexpr unsigned FBX = 0x80000, FBZ = 0x100000, FBY = 0x200000;

// Index bookkeeping of one level, passed to kernels by value.
struct Geo {
    int nox, noy, noz;
    int elx, ely, elz;
    int nno, nel;
    int pitch;     // size of the colour-ordered index space (whole 64-node tiles)
    int cs[9];     // first colour-ordered index of colour c (64-aligned), cs[8] = pitch
    int cn[8];     // nodes of colour c; indices cs[c]+cn[c] .. cs[c+1]-1 are padding
    int cz[8], cx[8]; // nodes of colour c along z and x
};

// A level cut into boxes of at most BOX nodes a side for the box smoother
// of the FP32 cycles: s[a][t] .. s[a][t+1]-1 are the node indices of the
// t-th segment along axis a (0 = z, 1 = x, 2 = y), n[a] segments.  A box of
// at most 8 nodes a side holds at most 4x4x4 = 64 nodes of each colour.
constexpr int BOX = 8;
constexpr int BOX_MAXSEG = 40;
struct BoxPart {
    int n[3];
    unsigned short s[3][BOX_MAXSEG + 1];
};

// arrival counter and XCD mask of a level's one-XCD smoother launches
// (k_gs_xcd); zeroed once at set-up, never reset by the launches
struct XcdBar {
    unsigned long long count;
    unsigned mask;
    unsigned pad;
};

struct Level {
    Geo g;
    // operator
    float *K = nullptr;      // [pitch/64][243][64]: 64-node tiles, colour order
    uint16_t *Kh = nullptr;  // FP16 bits of BI_row * K, same tiling (mixed = 2, fine levels only)
    double *BI = nullptr;    // [3 nno]
    double *BPI = nullptr;   // [nel]
    float *EG = nullptr;     // [nel][24]
    float *EVI = nullptr;    // [nel][8]
    // geometry
    float *GNX = nullptr;    // [nel][216]
    float *GDA = nullptr;    // [nel][11]
    float *ECO = nullptr;    // [nel][13]
    float *TWW = nullptr;    // [nel][9]
    float *MASS = nullptr;   // [nno]
    unsigned *NODE = nullptr;// [nno]
    float *iw = nullptr;     // interpolation weights n1,n2 per fine index: x[2*nox], z[2*noz], y[2*noy]
    float *VN = nullptr;     // [nno] nodal viscosity scratch
    // multigrid work vectors [3 nno]
    double *vel = nullptr, *res = nullptr, *fl = nullptr, *del = nullptr, *AU = nullptr;
    // coarsest level: its smoother call as a matrix [3nno][3nno], row-major
    double *Mc = nullptr;
    // level above the coarsest: its smoother call as d = Ma din + Mb F
    // ([3nno][3nno] each, row-major), for affine_sweeps sweeps; scratch
    // vectors for the result in both precisions
    double *Ma = nullptr, *Mb = nullptr;
    float *Maf = nullptr, *Mbf = nullptr;   // FP32 copies for the FP32 cycles (mixed forms)
    int affine_sweeps = 0;
    double *tmp = nullptr;
    float *tmpf = nullptr;
    // FP32 copies for the mixed-precision cycle
    float *velf = nullptr, *resf = nullptr, *flf = nullptr, *delf = nullptr, *AUf = nullptr;
    float *BIf = nullptr;
    // one-XCD smoother state (levels gs_xcd_ok accepts)
    XcdBar *xbar = nullptr;
    // box smoother of the FP32 cycles: the partition, and the cycle's
    // stiffness (Kh where the level has it, else K) re-tiled box by box,
    // per (box, colour) and y-plane each node's 81 coefficients in 16-byte
    // groups (k_box_tiles)
    BoxPart bp{};
    int nbox = 0;
    void *Kbox = nullptr;
    bool Kbox_half = false;
};

struct Prof {
    int kernel = 0;          // 0 off, 1 gs_colour, 2 matvec
    int level = -1;          // >= 0: only launches on this level
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<double> bytes;
    long launches = 0;
    double total_ms = 0.0, total_bytes = 0.0;
};

}  // namespace chip

struct chip_ctx {
    chip_problem p{};
    std::vector<float> xc, yc, zc;
    int levmin = 0, levmax = 0;
    chip::Level L[chip::MAXL];
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // finest-level state
    double *U = nullptr, *P = nullptr, *F = nullptr;
    float *T = nullptr, *Tdot = nullptr, *V = nullptr, *buoy = nullptr;
    float *T1 = nullptr, *Tdot1 = nullptr;
    double *Eres = nullptr;  // [nel][8] element residual / element force scratch
    float *TB = nullptr;     // [nno] boundary temperature (TB[3])
    float *expansivity = nullptr, *diffusivity = nullptr; // [noz]
    float *xc_d = nullptr, *yc_d = nullptr;               // finest coordinates per axis
    int *mat = nullptr;      // [nel] material group (1-based ids)
    float *H = nullptr;      // [noz] horizontal average
    float *hf = nullptr;     // [nno] nodal heat flux
    float *Have = nullptr;   // [3 noz] layer averages of T, |V|^2, viscosity (chip_averages)
    float *ave_scratch = nullptr;  // [nno]
    // Uzawa vectors
    double *r0 = nullptr, *r1 = nullptr, *r2 = nullptr, *z0 = nullptr, *z1 = nullptr;
    double *s1 = nullptr, *s2 = nullptr, *Ah = nullptr, *u1 = nullptr;
    // inner solve vectors
    double *sr = nullptr, *sd = nullptr;
    // reductions
    double *partial = nullptr;   // [4 * 1024]
    unsigned *counter = nullptr; // [1]
    double *dots = nullptr;      // [32] results, read by later kernels
    double *h_dots = nullptr;    // pinned host mirror
    unsigned *minmax = nullptr;  // [4] float bits for min/max reductions
    unsigned *h_minmax = nullptr;
    float diff_timestep = 0.0f;
    float elapsed_time = 0.0f;
    int solution_cycles = 0;
    bool built = false;      // operators constructed at least once
    bool use_graphs = true;  // replay multi_grid as a captured HIP graph
    bool small_levels = true;  // single-workgroup smoother on small levels
    bool gs_wg = true;         // ... including levels whose stiffness is read through L2
    bool xcd_gs = false;       // whole smoother calls of mid-size levels on one XCD (k_gs_xcd)
    // FP32 cycles: the levels between the exact (or affine) level and the
    // finest smoothed by boxes, one launch per sweep (option "box_gs")
    bool box_gs = true;
    // exact solve on the FP32 cycles' coarsest level (mixed forms, option
    // "coarse_direct"): the level (-1: none), its stiffness as a dense FP64
    // matrix factored and inverted in place (rocSOLVER, column-major), the
    // inverse (row-major; applied in FP16 below), the factorization's status
    int dlev = -1;
    bool direct_on = true;
    bool direct_ok = false;
    double *dA = nullptr;
    float *dAinv = nullptr;    // FP32 work buffer of the conversion
    // the inverse the cycles apply: FP16 rows scaled by their largest entry
    // (dAh, dAs), and the side stream's copies of both, copied in when taken up
    uint16_t *dAh = nullptr, *dAh_stage = nullptr;
    float *dAs = nullptr, *dAs_stage = nullptr;
    int *dinfo = nullptr;      // [2] potrf's and potri's status
    void *rb = nullptr;        // rocblas_handle
    // after the first, each factorization runs on a side stream during the
    // step it was started in and is taken up (converted to the FP32
    // inverse) at the next rebuild; until then the cycles use the previous
    // inverse
    hipStream_t side = nullptr;
    hipEvent_t ev_k = nullptr, ev_fact = nullptr;
    bool fact_pending = false;
    bool direct_async = true;  // option "direct_async": 0 factors every time on the work stream
    unsigned *xcd_ids = nullptr;   // [16] set-up check of the XCD dispatch
    // learned coarse-level correction (chip_set_coarse_net): a 3-D CNN on the
    // coarsest level in place of its smoother call; BF16 weights
    // [layers][16][27][16], FP32 biases [layers][16], output scales [3],
    // per-node log10 viscosity
    int cnet_layers = 0;
    bool cnet_on = false;
    bool cnet_failed = false;   // a launch of the correction failed while enqueueing
    int cnet_fallbacks = 0;     // inner solves restarted without it (no residual reduction)
    bool cnet_calibrated = false;  // the first solve with it compared it with the plain multigrid
    uint16_t *cnet_w = nullptr;
    float *cnet_b = nullptr, *cnet_ys = nullptr, *cnet_logeta = nullptr;
    double *cnet_keep = nullptr;   // [3 nno] the calibration's plain result
    // warm start of the Uzawa solve (option "warm_start"): 0 the previous
    // solution as it is (the reference); 1 extrapolated along the sequence
    // of solves from the last two or three (warm_start_guess); 2 the same
    // plus the learned correction (chip_set_warm_net)
    int warm_start = 0;
    int hist = 0;                 // solutions kept since the option was set
    double *Uprev = nullptr, *Pprev = nullptr;    // the two solutions before the current one
    double *Uprev2 = nullptr, *Pprev2 = nullptr;
    double *D2u = nullptr, *D2p = nullptr;        // their second differences
    float *Th0 = nullptr, *Th1 = nullptr;         // T at the last two solves
    float *wn_d2pn = nullptr;
    // learned correction of the extrapolated guess (chip_set_warm_net): the
    // U-Net's BF16 weights and FP32 biases, its input [nno][16] and output
    // [nno][16], its workspace, and the finest-level scratch of the features
    uint16_t *wn_w = nullptr, *wn_x = nullptr;
    float *wn_b = nullptr, *wn_out = nullptr;
    void *wn_ws = nullptr;
    float *wn_dT = nullptr, *wn_pn = nullptr, *wn_dpn = nullptr, *wn_logeta = nullptr;
    double *Ue = nullptr, *Pe = nullptr;        // the extrapolated guess of the current solve
    bool wn_set = false;
    // training samples (chip_warm_record): the features and the target of
    // each solve that has two earlier ones, into caller device buffers
    uint16_t *rec_x = nullptr;
    float *rec_y = nullptr;
    bool feat_pending = false;   // this solve's features were formed; its target follows the solve
    double wscale[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // s_u, s_p, s_dT, s_2u, s_2p, (unused), hist, formed
    bool print_convergence = false;  // per-iteration lines on stderr, as E->control.print_convergence
    hipGraphExec_t mg_graph = nullptr;
    chip::Prof prof;
    long dev_bytes = 0;
    std::vector<void *> allocs;
};

namespace chip {
// setup (chip_setup.cpp)
int build_mesh(chip_ctx *c);
// device buffers of the learned coarse-level correction (once per context)
constexpr int CNET_MAX_LAYERS = 8;
int alloc_coarse_net(chip_ctx *c);
// history buffers of the warm start (once per context)
int alloc_warm_start(chip_ctx *c);
// the learned correction's buffers (once per context)
int alloc_warm_net(chip_ctx *c);
// errors
void set_error(const std::string &s);
}  // namespace chip

#define CHIP_CHECK(expr)                                                          \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            chip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
            return -1;                                                            \
        }                                                                         \
    } while (0)
