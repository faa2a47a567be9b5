// xyce_amd_master.hpp -- host-vector Master over the C ABI (C++17,
// header-only; needs xyce_amd.h and libxyce_amd.so, nothing else).
//
// Mirrors the Master virtuals a Xyce device type implements
// (src/DeviceModelPKG/Core/include/N_DEV_Device.h, default loops
// N_DEV_DeviceMaster.h:643-725; per-model overrides N_DEV_Diode.C:2051-2171,
// N_DEV_MOSFET1.C:4069-4688, N_DEV_BJT.C:3891-4366,
// N_DEV_MOSFET_B4.C:16180-17268) for a caller whose circuit vectors are host
// arrays, as in a plain Xyce build:
//
//   bool updateState(double* solVec, double* staVec, double* stoVec);
//   bool updateSecondaryState(double* staDerivVec, double* stoVec);
//   bool loadDAEVectors(double* solVec, double* fVec, double* qVec,
//                       double* bVec, double* storeLeadF, double* storeLeadQ);
//   bool loadDAEMatrices(CsrMatrix& dFdx, CsrMatrix& dQdx);
//
// The other vectors a load reads (current/last solution, state and store,
// the limiter vectors) come through ExternData, named after
// N_DEV_ExternData.h, as the reference's instances reach them.
//
// Each call copies the vectors it reads to the device, runs the load there,
// and copies back the vectors it writes.  Output vectors go up holding what
// the host already has (the other device types' contributions: DeviceMgr
// zeroes them, N_DEV_DeviceMgr.C:2165-2174, then every Master adds), and the
// device adds this type's contributions onto those values in the reference's
// instance order, so the host sees the reference's sums.  Store and state
// vectors go up whole and come back whole: the slots of other device types
// are returned unchanged.
//
// Setup errors throw std::runtime_error (the reference throws at setup,
// N_DEV_DeviceMaster.C:75-100); load calls return false, as the reference's
// Master functions do, with the message in lastError().
#ifndef XYCE_AMD_MASTER_HPP
#define XYCE_AMD_MASTER_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "xyce_amd.h"

namespace xyce_amd {

// The raw vector pointers of N_DEV_ExternData.h that the loads read besides
// their arguments (host arrays; nullptr where the analysis has none).
struct ExternData {
  double* currSolVectorRawPtr = nullptr;
  double* currStaVectorRawPtr = nullptr;
  double* lastStaVectorRawPtr = nullptr;
  double* currStoVectorRawPtr = nullptr;
  double* dFdxdVpVectorRawPtr = nullptr;
  double* dQdxdVpVectorRawPtr = nullptr;
  double* flagSolVectorRawPtr = nullptr;   // flagSolVectorPtr's values (inputOP)
};

// A CSR matrix as the loads see it: the structure registered at setup and a
// host array of values (the role of N_LAS_Matrix in loadDAEMatrices).
struct CsrMatrix {
  int n_rows = 0;
  const int32_t* row_ptr = nullptr;
  const int32_t* col_idx = nullptr;
  double* values = nullptr;
};

class Master {
 public:
  // One Master per device type (XA_DIODE, XA_MOSFET1, XA_BJT, XA_BSIM4).
  explicit Master(int model_type, const xa_device_options* options = nullptr, int device = 0) {
    check(xa_ctx_create(device, &ctx_), "xa_ctx_create");
    xa_device_options o;
    xa_default_options(&o);
    if (options) o = *options;
    if (xa_group_create(ctx_, model_type, &o, &g_) != XA_OK) {
      const std::string msg = xa_last_error();
      xa_ctx_destroy(ctx_);
      throw std::runtime_error("xa_group_create: " + msg);
    }
    xa_default_solver_state(&ss_);
  }
  ~Master() {
    for (void* p : owned_) xa_device_free(ctx_, p);
    xa_group_destroy(g_);
    xa_ctx_destroy(ctx_);
  }
  Master(const Master&) = delete;
  Master& operator=(const Master&) = delete;

  // ---- setup, in the ABI's order ----
  int addModel(int level, const std::vector<std::string>& names, const std::vector<double>& values) {
    std::vector<const char*> cn = cstrs(names);
    int idx = -1;
    check(xa_group_add_model(g_, level, (int)cn.size(), cn.data(), values.data(), &idx),
          "xa_group_add_model");
    return idx;
  }
  // columns[p][i]: instance parameter names[p] of instance i
  void addInstances(int n, const int* model_index, const std::vector<std::string>& names,
                    const std::vector<const double*>& columns) {
    std::vector<const char*> cn = cstrs(names);
    check(xa_group_add_instances(g_, n, model_index, (int)cn.size(), cn.data(), columns.data()),
          "xa_group_add_instances");
    n_ = n;
  }
  void requestLeadCurrent(const int32_t* flags) {
    check(xa_group_request_lead_current(g_, flags), "xa_group_request_lead_current");
  }
  void setup() { check(xa_group_setup(g_), "xa_group_setup"); }
  // per-instance external nodes, internal nodes, state and store variables
  void counts(int* n_ext, int* n_int, int* n_sta, int* n_sto) const {
    check(xa_group_counts(g_, n_ext, n_int, n_sta, n_sto), "xa_group_counts");
  }
  void registerLIDs(const int32_t* node_lids, const int32_t* sta_lids, const int32_t* sto_lids) {
    check(xa_group_register_lids(g_, node_lids, sta_lids, sto_lids), "xa_group_register_lids");
  }
  // jacobianStamp(): (row, col) pairs after node collapse
  void jacPairs(std::vector<int32_t>& rows, std::vector<int32_t>& cols) const {
    int64_t cnt = 0;
    check(xa_group_jac_pairs(g_, &cnt, nullptr, nullptr), "xa_group_jac_pairs");
    rows.resize((size_t)cnt);
    cols.resize((size_t)cnt);
    check(xa_group_jac_pairs(g_, &cnt, rows.data(), cols.data()), "xa_group_jac_pairs");
  }
  void registerCSR(int n_rows, const int32_t* row_ptr, const int32_t* col_idx) {
    check(xa_group_register_csr(g_, n_rows, row_ptr, col_idx), "xa_group_register_csr");
    nnz_ = n_rows > 0 ? (size_t)row_ptr[n_rows] : 0;
  }
  // Vector lengths of the circuit: solution (and F, Q, B, limiter vectors),
  // state and store (and lead-current) vectors.
  void finalize(size_t solution_size, size_t state_size, size_t store_size) {
    check(xa_group_finalize(g_), "xa_group_finalize");
    n_sol_ = solution_size;
    n_sta_ = state_size;
    n_sto_ = store_size;
    double** fields[] = {&d_.nextSol_, &d_.currSol_, &d_.flagSol_, &d_.f_, &d_.q_, &d_.b_,
                         &d_.dFdxdVp_, &d_.dQdxdVp_};
    for (double** f : fields) *f = alloc(n_sol_);
    double** sta[] = {&d_.nextSta_, &d_.currSta_, &d_.lastSta_, &d_.nextStaDeriv_};
    for (double** f : sta) *f = alloc(n_sta_);
    double** sto[] = {&d_.nextSto_, &d_.currSto_, &d_.leadF_, &d_.leadQ_};
    for (double** f : sto) *f = alloc(n_sto_);
    d_.dFdx_ = alloc(nnz_);
    d_.dQdx_ = alloc(nnz_);
  }

  // SolverState flags of the current load (N_DEV_SolverState.h:93-190).
  xa_solver_state& solverState() { return ss_; }
  ExternData& externData() { return ext_; }
  std::string lastError() const { return err_; }

  // ---- the Master virtuals ----
  bool updateState(double* solVec, double* staVec, double* stoVec) {
    err_.clear();
    xa_vectors v = vectors();
    ok_ = put(v.nextSol, solVec, n_sol_) && put(v.currSol, ext_.currSolVectorRawPtr, n_sol_) &&
          put(v.flagSol, ext_.flagSolVectorRawPtr, n_sol_) && put(v.nextSta, staVec, n_sta_) &&
          put(v.currSta, ext_.currStaVectorRawPtr, n_sta_) &&
          put(v.lastSta, ext_.lastStaVectorRawPtr, n_sta_) && put(v.nextSto, stoVec, n_sto_) &&
          put(v.currSto, ext_.currStoVectorRawPtr, n_sto_);
    if (!ok_) return false;
    // a vector the host does not have reaches the kernels as null (never as
    // the device buffer's values from an earlier call); the loads reject a
    // null vector where the solver state needs it
    if (!ext_.flagSolVectorRawPtr) v.flagSol = nullptr;
    if (!ext_.currSolVectorRawPtr) v.currSol = nullptr;
    if (!ext_.currStaVectorRawPtr) v.currSta = nullptr;
    if (!ext_.lastStaVectorRawPtr) v.lastSta = nullptr;
    if (!ext_.currStoVectorRawPtr) v.currSto = nullptr;
    if (!run(xa_update_state(g_, &ss_, &v), "xa_update_state")) return false;
    // the state slots this type writes, and the initial-transient copies
    // into currSta / lastSta (N_DEV_MOSFET_B4.C:16254-16289, N_DEV_BJT.C:3923-3933)
    return get(staVec, v.nextSta, n_sta_) && get(stoVec, v.nextSto, n_sto_) &&
           get(ext_.currStaVectorRawPtr, v.currSta, n_sta_) &&
           get(ext_.lastStaVectorRawPtr, v.lastSta, n_sta_);
  }
  bool updateSecondaryState(double* staDerivVec, double* stoVec) {
    err_.clear();
    xa_vectors v = vectors();
    if (!put(v.nextStaDeriv, staDerivVec, n_sta_) || !put(v.nextSto, stoVec, n_sto_)) return false;
    if (!run(xa_update_secondary_state(g_, &ss_, &v), "xa_update_secondary_state")) return false;
    return get(stoVec, v.nextSto, n_sto_);
  }
  bool loadDAEVectors(double* solVec, double* fVec, double* qVec, double* bVec,
                      double* storeLeadF, double* storeLeadQ) {
    err_.clear();
    xa_vectors v = vectors();
    if (!(put(v.nextSol, solVec, n_sol_) && put(v.f, fVec, n_sol_) && put(v.q, qVec, n_sol_) &&
          put(v.b, bVec, n_sol_) && put(v.dFdxdVp, ext_.dFdxdVpVectorRawPtr, n_sol_) &&
          put(v.dQdxdVp, ext_.dQdxdVpVectorRawPtr, n_sol_) &&
          put(v.storeLeadF, storeLeadF, n_sto_) && put(v.storeLeadQ, storeLeadQ, n_sto_)))
      return false;
    if (!fVec) v.f = nullptr;
    if (!qVec) v.q = nullptr;
    if (!bVec) v.b = nullptr;
    if (!ext_.dFdxdVpVectorRawPtr) v.dFdxdVp = nullptr;
    if (!ext_.dQdxdVpVectorRawPtr) v.dQdxdVp = nullptr;
    if (!storeLeadF) v.storeLeadF = nullptr;
    if (!storeLeadQ) v.storeLeadQ = nullptr;
    if (!run(xa_load_vectors(g_, &ss_, &v), "xa_load_vectors")) return false;
    return get(fVec, v.f, n_sol_) && get(qVec, v.q, n_sol_) && get(bVec, v.b, n_sol_) &&
           get(ext_.dFdxdVpVectorRawPtr, v.dFdxdVp, n_sol_) &&
           get(ext_.dQdxdVpVectorRawPtr, v.dQdxdVp, n_sol_) &&
           get(storeLeadF, v.storeLeadF, n_sto_) && get(storeLeadQ, v.storeLeadQ, n_sto_);
  }
  bool loadDAEMatrices(CsrMatrix& dFdx, CsrMatrix& dQdx) {
    err_.clear();
    xa_vectors v = vectors();
    if (!put(v.dFdx, dFdx.values, nnz_) || !put(v.dQdx, dQdx.values, nnz_)) return false;
    if (!dFdx.values) v.dFdx = nullptr;
    if (!dQdx.values) v.dQdx = nullptr;
    if (!run(xa_load_matrices(g_, &ss_, &v), "xa_load_matrices")) return false;
    return get(dFdx.values, v.dFdx, nnz_) && get(dQdx.values, v.dQdx, nnz_);
  }

  // Instances whose evaluation reported failure in the last update.
  int64_t failures() const {
    int64_t n = 0;
    check(xa_group_failures(g_, &n), "xa_group_failures");
    return n;
  }
  int numInstances() const { return n_; }
  xa_ctx context() const { return ctx_; }
  xa_group group() const { return g_; }

 private:
  struct DeviceVectors {
    double *nextSol_ = nullptr, *currSol_ = nullptr, *flagSol_ = nullptr, *nextSta_ = nullptr,
           *currSta_ = nullptr, *lastSta_ = nullptr, *nextStaDeriv_ = nullptr,
           *nextSto_ = nullptr, *currSto_ = nullptr, *f_ = nullptr, *q_ = nullptr, *b_ = nullptr,
           *dFdxdVp_ = nullptr, *dQdxdVp_ = nullptr, *leadF_ = nullptr, *leadQ_ = nullptr,
           *dFdx_ = nullptr, *dQdx_ = nullptr;
  };

  static void check(int rc, const char* what) {
    if (rc != XA_OK) throw std::runtime_error(std::string(what) + ": " + xa_last_error());
  }
  static std::vector<const char*> cstrs(const std::vector<std::string>& names) {
    std::vector<const char*> out;
    for (const std::string& s : names) out.push_back(s.c_str());
    return out;
  }
  double* alloc(size_t n) {
    void* p = nullptr;
    check(xa_device_alloc(ctx_, (n > 0 ? n : 1) * sizeof(double), &p), "xa_device_alloc");
    owned_.push_back(p);
    return static_cast<double*>(p);
  }
  xa_vectors vectors() const {
    xa_vectors v{};
    v.nextSol = d_.nextSol_;
    v.currSol = d_.currSol_;
    v.flagSol = d_.flagSol_;
    v.nextSta = d_.nextSta_;
    v.currSta = d_.currSta_;
    v.lastSta = d_.lastSta_;
    v.nextStaDeriv = d_.nextStaDeriv_;
    v.nextSto = d_.nextSto_;
    v.currSto = d_.currSto_;
    v.f = d_.f_;
    v.q = d_.q_;
    v.b = d_.b_;
    v.dFdxdVp = d_.dFdxdVp_;
    v.dQdxdVp = d_.dQdxdVp_;
    v.storeLeadF = d_.leadF_;
    v.storeLeadQ = d_.leadQ_;
    v.dFdx = d_.dFdx_;
    v.dQdx = d_.dQdx_;
    return v;
  }
  bool fail(const char* what) {
    err_ = std::string(what) + ": " + xa_last_error();
    return false;
  }
  // host -> device when the host vector exists
  bool put(const double* dev, const double* host, size_t n) {
    if (!host || n == 0) return true;
    return xa_copy_to_device(ctx_, const_cast<double*>(dev), host, n * sizeof(double)) == XA_OK
               ? true
               : fail("xa_copy_to_device");
  }
  bool get(double* host, const double* dev, size_t n) {
    if (!host || !dev || n == 0) return true;
    return xa_copy_to_host(ctx_, host, dev, n * sizeof(double)) == XA_OK ? true
                                                                         : fail("xa_copy_to_host");
  }
  bool run(int rc, const char* what) { return rc == XA_OK ? true : fail(what); }

  xa_ctx ctx_ = nullptr;
  xa_group g_ = nullptr;
  xa_solver_state ss_{};
  ExternData ext_{};
  DeviceVectors d_{};
  std::vector<void*> owned_;
  std::string err_;
  size_t n_sol_ = 0, n_sta_ = 0, n_sto_ = 0, nnz_ = 0;
  int n_ = 0;
  bool ok_ = true;
};

}  // namespace xyce_amd

#endif  // XYCE_AMD_MASTER_HPP
