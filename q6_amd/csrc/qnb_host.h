// qnb_host.h -- host-side pieces shared by the single-system and batch ABIs.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/qnb.h"
#include "qnb_dev.h"

namespace qnb {

int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) return ::qnb::fail(QNB_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

// device buffers owned by one handle, freed with it
struct DevAlloc {
  std::vector<void*> ptrs;
  ~DevAlloc() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T** out, size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return fail(QNB_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
    ptrs.push_back(p);
    *out = (T*)p;
    return 0;
  }
  template <typename T>
  int upload(const T** out, const T* src, size_t n) {
    T* p;
    int rc = alloc(&p, n);
    if (rc) return rc;
    if (n && src && hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return fail(QNB_EDEVICE, "upload failed");
    *out = p;
    return 0;
  }
  void release(void* p) {
    auto it = std::find(ptrs.begin(), ptrs.end(), p);
    if (it != ptrs.end()) {
      (void)hipFree(p);
      ptrs.erase(it);
    }
  }
};

// host copies of the topology kept for list export
struct HostTopo {
  int ntypes = 0, ivdw_rule = 1;
  std::vector<int> gfirst, glast, gswitch, cgpatom, iqatom, iac, ljcod, lx_off, lx_j, iqseq;
  std::vector<uint8_t> excl, qconn3, lx_kind;
  std::vector<uint32_t> exmask, m14mask;
  std::vector<double> crg, avdw, bvdw;
  std::vector<int> qq_iq, qq_jq, qqp_iq, qqp_j;  // state-0 static lists
};

int validate_system(const qnb_system* s, const qnb_params* p);
// sizes, cutoffs, lambda, SCAAS constants of one instance
void fill_params(DevSys& S, const qnb_system* s, const qnb_params* p);
// derived tables and device copies of the topology (shareable by instances)
int topo_build(DevSys& S, const qnb_system* s, HostTopo& T, DevAlloc& M);
// host twin of the device solute-solute filter (pp_param), for list export
bool host_pp_param(const HostTopo& T, int i, int j);
// per-instance device state (coordinates, gradient, lists, LRF, energies)
int instance_alloc(DevSys& S, DevAlloc& M, const qnb_params* p, int nslots);
// group contributions (qnb_exc.hip): the pair lists of each energy array
// and state, built on the host from the system's Q tables
struct ExcHost {
  int narr = 0, ns = 0;
  std::vector<int> off, type;
  std::vector<ExcEnt> ent;
};
int exc_build(const qnb_system* s, int narr, const uint8_t* mask, const int32_t* type, ExcHost& X);
int exc_upload(const ExcHost& X, ExcDev& D, DevAlloc& M);
void launch_exc_eval(hipStream_t st, const ExcDev& D, const double* x, const double* eq_in, double* out);
// free-energy analysis (qnb_fep.hip): one window is either a batch
// instance's recorded frames (hist) or state energies the caller supplies
// (V: EQ(s)%total per frame; rkl: the first off-diagonal's distance per frame
// or none)
struct FepWin {
  const double* hist;  // [nf][QNB_FRAME_LEN] or null
  const double* V;     // [nf][ns] or null
  const double* rkl;   // [nf] or null
  int nf;              // frames
  int noffd;           // off-diagonals recorded per frame
  double lam[QNB_MAX_STATES];
};
struct FepPar {
  double kT;
  int nskip, nbins, nmin, ns;
  double alpha[QNB_MAX_STATES], coeff[QNB_MAX_STATES];
  double A, mu, eta, r0;
};
int fep_params(const qnb_fep_params* p, int ns, FepPar* P);
// runs the analysis on stream st for windows whose data sit in device memory
// (hw: host copies of the window records) and copies the results to out
int fep_run(hipStream_t st, const std::vector<FepWin>& hw, const FepPar& P, qnb_fep_out* out);

// adjoint of the Q-atom energies: tangent upload and launch over frames
int adjoint_upload(DevAlloc& M, const DevSys& S, const qnb_qtangent* t, AdjDev& A);
void launch_adjoint(hipStream_t st, int prec, int ns_max, long waves, int B, const DevSys* dsys, const AdjDev* dadj);
// Q-atom bonded terms: validation and device copies (buffers recorded in
// `owned` so a later set can release them), and the per-instance launch
int qbonded_build(const qnb_qbonded* q, int nstates, int natom, const double* lambda, DevAlloc& M,
                  std::vector<void*>& owned, QbDev& out);
void qbonded_free(DevAlloc& M, std::vector<void*>& owned);
void launch_qbonded(hipStream_t st, int B, const DevSys* dsys, const QbDev* dqb);

}  // namespace qnb
