// fm_internal.h -- private structures of the MI355X FEMuS hot-path library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/femus_hip.h"
#include "fm_mesh.h"
#include "egather_host.h"

void fm_set_error(const std::string &msg);
unsigned fm_error_count();  // fm_set_error calls so far on this thread
// invalid-argument return of an entry point: names the function, so fm_last_error never reports an older failure
inline int fm_einval(const char *where) {
  fm_set_error(std::string(where) + ": invalid argument");
  return FM_EINVAL;
}
// propagates a failing internal call; names the call when it set no message of its own
#define FM_CHECK(x)                                                  \
  do {                                                               \
    const unsigned e0_ = fm_error_count();                           \
    const int rc_ = (x);                                             \
    if (rc_ != 0) {                                                  \
      if (fm_error_count() == e0_) fm_set_error(#x);                 \
      return rc_ < 0 ? rc_ : FM_EHIP;                                \
    }                                                                \
  } while (0)

#define FM_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fm_set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return FM_EHIP;                                                          \
    }                                                                          \
  } while (0)

// Sliced ELLPACK copy (SELL-C-sigma, C = 64 = one wavefront) in the level's
// internal DoF order (see Level::ord): rows sorted by length inside windows of
// 4096, slice s holds internal rows perm[64s .. 64s+63] with width sl_w[s];
// entry k of lane l sits at slot sl_off[s] + 64k + l (column-major inside the
// slice, so a wavefront's loads are contiguous).  Inside each row the entries
// are sorted by internal column.  Columns are stored as offsets from a
// per-slot-column base, base[sl_off[s]/64 + k], in the narrowest of four classes per slice:
//   2-bit lane-predicted (every column of position k is base + l + d, d in [0, 3], base = the smallest
//          column - lane): d of position k, lane l in bits 2 (k%16) of c2[sl_c2[s] + 64 (k/16) + l];
// otherwise base is the smallest column among the 64 lanes at position k and
//   8-bit  (every position spans <= 255 columns): offset of position k, lane l
//          at byte c8[sl_c8[s] + 256 (k/4) + 4 l + k%4] (four positions per
//          32-bit word, the 64 words of a group contiguous);
//   16-bit (spans <= 65535): c16[sl_c16[s] + 128 (k/2) + 2 l + k%2];
//   32-bit absolute columns otherwise: c32[sl_c32[s] + 64 k + l].
// The offsets of the other classes are -1.  Padding decodes to one of the
// position's real columns with value 0; pos / rbase / rslot tie every slot to its CSR entry.
struct DevSELL {
  int64_t nsl = 0, nslot = 0, n32 = 0, n8 = 0, n16 = 0, n2 = 0;
  int64_t *sl_off = nullptr, *sl_c32 = nullptr, *sl_c8 = nullptr, *sl_c16 = nullptr, *sl_c2 = nullptr;
  unsigned *c2 = nullptr;
  int *sl_w = nullptr, *perm = nullptr, *base = nullptr, *c32 = nullptr;
  // slot -> CSR entry: slot (s, k, l) of internal row i = perm[64 s + l] holds CSR entry rbase[i] + pos[slot]
  // (pos 0xFFFF: padding); rslot[i]: slot of position 0 of internal row i
  unsigned short *pos = nullptr;
  int64_t *rbase = nullptr, *rslot = nullptr;
  bool borrowed = false;  // layout arrays belong to another matrix of the same pattern (only v / vf are owned)
  unsigned short *c16 = nullptr;
  unsigned char *c8 = nullptr;
  double *v = nullptr;
  float *vf = nullptr;
  std::vector<int64_t> h_gslice;  // first slice of each row group (+ end)
};

// Device CSR (int64 row pointers, int32 columns); values in FP64 and FP32.
struct DevCSR {
  int64_t nr = 0, nc = 0, nnz = 0;
  int64_t *rp = nullptr;
  int *ci = nullptr;
  double *v = nullptr;
  float *vf = nullptr;
  // host copies (pattern + values as built)
  std::vector<int64_t> h_rp;
  std::vector<int> h_ci;
  std::vector<double> h_v;
  DevSELL sell;
};

// Element-buffer assembly of one level: the element matrices go to a buffer E[e][row][col] (NL x NL doubles
// per element, row / column in the orientation of the target matrix), and a gather sums them into the SELL
// slots (the plan of egather_host.h, on the device).
struct EGather {
  bool built = false, unusable = false;
  int lg = 0, nl = 0, maxseg = 0, maxctr = 0, maxit = 0;  // per block: element rows, contributions, slots
  int64_t nblk = 0, nseg = 0, nctr = 0;
  int64_t *lseg = nullptr, *bbase = nullptr;
  int *seg = nullptr;
  unsigned char *cnt = nullptr;
  unsigned short *cst = nullptr, *cidx = nullptr;
  size_t bytes = 0;
};

// entries past the end of the element -> slot maps: the 3D patch inversion reads its 89 x 89 map block as a
// 96 x 96 tile (up to 95 * 89 + 95 = 8550 entries from a patch's first one, 629 past the last patch's block)
constexpr size_t FM_ELMAP_PAD = 1024;

struct Level {
  int64_t n = 0;           // _Dim[L]
  int nel = 0;
  // internal DoF order of the level's multigrid vectors and SELL operators:
  // ord[i] = reference DoF (MGSolDA::_node_dof numbering) at internal position i,
  // rank[dof] = internal position of a reference DoF
  std::vector<int> h_ord, h_rank;
  int *ord = nullptr, *rank = nullptr;
  // row groups of equal colour signature (coupled systems): internal row ranges and signatures;
  // per colour c the SELL slices of the groups it touches: cslist[cs_off[c] .. cs_off[c+1])
  std::vector<int64_t> h_goff;
  std::vector<int> h_gmask;
  int *cslist = nullptr;
  std::vector<int64_t> cs_off;
  DevCSR A, R, P;          // A[L]; R: coarse L <- fine L+1; P: fine L <- coarse L-1
  // FP64 vectors
  double *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr, *d = nullptr, *d2 = nullptr, *dinv = nullptr;
  // FP32 vectors
  float *xf = nullptr, *bf = nullptr, *rf = nullptr, *tf = nullptr, *df = nullptr, *d2f = nullptr, *dinvf = nullptr;
  double lam = 0.0;        // lambda_max(D^-1 A) estimate
  // assembly support
  int *conn = nullptr;     // nel * n2 node ids
  int *eldof = nullptr;    // nel * nloc dofs (reference numbering)
  int *eldofi = nullptr;   // nel * nloc dofs (internal numbering)
  // element map of A (the assembly writes the SELL values directly): entry (a, b) of element e sits at SELL
  // slot elslot[e][a] + 64 elpos[e][a][b] (elpos 0xFFFF: no entry); both arrays carry FM_ELMAP_PAD entries
  // more (0xFFFF / -1), read by the Vanka patch inversion's padded tiles
  unsigned short *elpos = nullptr;  // nel * nloc * nloc positions within the row
  int64_t *elslot = nullptr;        // nel * nloc: slot of position 0 of the row of each element DoF (-1: none)
  int64_t *dslot = nullptr;         // n: SELL slot of the diagonal of internal row i
  unsigned char *bmask = nullptr;  // nel: bit s set if side s on the boundary
  int ncolor = 0;
  std::vector<int> color_off; // element ranges per colour (elements reordered by colour)
  int *color_elems = nullptr; // element ids grouped by colour
  double *coef = nullptr;  // per-element k or nu (nullptr -> default)
  // coarse direct solve (level 0)
  double *Ainv = nullptr;
  float *Ainvf = nullptr;
  // Vanka local inverses (nel * nloc * nloc, transposed; FP64 and its FP32 copy)
  double *vinv = nullptr;
  float *vinvf = nullptr;
  double *vwt = nullptr;   // 1 / multiplicity per dof
  double *vone = nullptr;  // all ones (coloured Vanka weights)
  float *vwtf = nullptr;
  // norms for the reference correction rule
  double *nrm = nullptr;   // device scalars
  // residual for transfers
  double *res = nullptr;
  // saved rhs / iterate while an outer Krylov method borrows the level's cycle vectors
  double *sb = nullptr, *sx = nullptr;
  EGather eg;              // element-buffer assembly (built at the first assembly that uses it)
};

// the element-buffer gather of a level built from its slot map and SELL layout (femus.cpp); 0, or -1 with err
// set when the blocks cannot be staged in max_lds bytes
int fm_build_egather(hipStream_t st, Level &l, int nl, size_t max_lds, EGather &g, std::string &err);
void fm_free_egather(EGather &g);

struct fm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double *scratch = nullptr;    // reduction scratch (device)
  double *h_scratch = nullptr;  // pinned host mirror
  int nblocks_red = 1024;
  int nsys = 0;                 // live systems created on this context
};

// a captured V-cycle (one per precision / option set / forward-or-adjoint)
struct GraphEntry {
  int prec, adj, xz, smoother, nu_pre, nu_post, correction, gamma, colored, cmult;
  int split;  // split-slice threshold of the level products at capture (the launch shapes it recorded)
  double lo, hi, omega;
  hipGraphExec_t exec;
};

// An operator handed in by the caller in the reference numbering (ReadMatrix / ReadRest / ReadProl): CSR with
// sorted columns; values only for the transfers
struct UserCsr {
  bool set = false;
  int64_t nc = 0;
  std::vector<int64_t> rp;
  std::vector<int> ci;
  std::vector<double> v;
};

struct FeTab;
struct fm_system {
  fm_ctx *ctx = nullptr;
  const fm_mesh *mesh = nullptr;
  int nq2 = 0, nq1 = 0, nloc = 0;
  std::vector<std::vector<int>> node_dof;   // per level, (nq2+nq1)*Ntop
  std::vector<int64_t> Dim;
  std::vector<int> bc0, bc1;                // top-level DoFs
  std::vector<UserCsr> u_pat, u_rst, u_prl;  // per level: caller's pattern of A, restriction, prolongation
  int *d_bc0 = nullptr, *d_bc1 = nullptr;   // device copies (top DoFs)
  int *d_dof_top = nullptr;                 // node_dof[top] on device
  double *xyz = nullptr;                    // device coords
  double *x_old = nullptr;                  // top-level x_old (device)
  double *vel = nullptr;                    // heat advection field [dim*Ntop] or nullptr
  double *xbc = nullptr;                    // Dirichlet data of the penalty RHS (nullptr: x_old)
  double *tnode = nullptr;                  // Boussinesq: nodal temperature [Ntop] (nullptr: no buoyancy)
  double buoy[3] = {0, 0, 0};
  std::vector<Level> lev;
  // transposed hierarchy for the discrete adjoint: A^T (same pattern as A),
  // restriction Prl[L+1]^T, prolongation Rst[L-1]^T, own work vectors
  std::vector<Level> alev;
  std::vector<char> a_assembled;
  bool adj_ready = false;
  std::vector<GraphEntry> graphs;
  int adj_smoother = -1;
  bool initialized = false;
  bool mg_ready = false;
  int mg_smoother = -1;
  int vanka_store = 3;  // precisions of the stored Vanka inverses: bit 0 FP64 (FP64 cycles), bit 1 FP32 (FP32 cycles)
  FeTab *d_tab = nullptr;  // reference-element tables on the device (built at the first fm_system_init)
  // flexible-GMRES workspace of the top level, kept from one solve to the next (fm_mg_release frees it):
  // basis V [(km + 1) kn], preconditioned vectors Z [km kn], Gram-Schmidt partial sums, the Hessenberg
  // state (coefficients, rotations, residual estimates) and a pinned mirror of the estimates with one event
  // per Arnoldi step
  double *kV = nullptr, *kZ = nullptr, *kpart = nullptr, *kH = nullptr, *kpin = nullptr;
  int64_t kn = 0;
  int km = 0, knb = 0;
  std::vector<hipEvent_t> kev;
  // element-matrix and element-rhs buffers of the element-buffer assembly (sized for the largest level that
  // uses it; shared by the levels and by the transposed assemblies)
  double *ebuf = nullptr, *febuf = nullptr;
  size_t ebuf_n = 0, febuf_n = 0;
};
void fm_krylov_free(fm_system *s);

// views of a device matrix for the kernel launchers (SELL copy when present)
#include "fm_kernels.h"
template <typename T> inline void fm_sell_fields(CsrView<T> &v, const DevSELL &q) {
  v.nsl = q.nsl; v.nslot = q.nslot; v.sl_off = q.sl_off; v.sl_w = q.sl_w; v.perm = q.perm; v.c16 = q.c16; v.base = q.base;
  v.sl_c32 = q.sl_c32; v.c32 = q.c32; v.sl_c8 = q.sl_c8; v.c8 = q.c8; v.sl_c16 = q.sl_c16; v.sl_c2 = q.sl_c2; v.c2 = q.c2;
}
template <typename T> CsrView<T> fm_view(const DevCSR &c);
template <> inline CsrView<double> fm_view<double>(const DevCSR &c) {
  CsrView<double> v;
  v.nr = c.nr; v.rp = c.rp; v.ci = c.ci; v.v = c.v;
  v.lpr = fmk::lanes_for(c.nnz, c.nr);
  fm_sell_fields(v, c.sell);
  v.sv = c.sell.v;
  return v;
}
template <> inline CsrView<float> fm_view<float>(const DevCSR &c) {
  CsrView<float> v;
  v.nr = c.nr; v.rp = c.rp; v.ci = c.ci; v.v = c.vf;
  v.lpr = fmk::lanes_for(c.nnz, c.nr);
  fm_sell_fields(v, c.sell);
  v.sv = c.sell.vf;
  return v;
}
// ro[i] = CSR row stored at internal row i; cr[j] = internal column of CSR column j (nullptr = identity)
int fm_build_sell(hipStream_t st, DevCSR &c, const int *ro, const int *cr, const std::vector<int64_t> *groups);
int fm_alloc_level_vectors(Level &l, bool top);
void fm_drop_graphs(fm_system *s);
void fm_free_level(Level &l);
int fm_sync_sell(hipStream_t st, DevCSR &c);
// A's values live in its SELL copy; this refreshes the CSR copy's FP64/FP32 values from it
int fm_csr_from_sell(hipStream_t st, DevCSR &c);
void fm_free_csr(DevCSR &c);
