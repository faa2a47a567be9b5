// xa_internal.h -- host-side structures shared by the runtime and the
// per-model back ends (not part of the public ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "xyce_amd.h"

namespace xa {

// Internal output kinds: the public six plus the two lead-current store
// vectors, which are assigned (not accumulated) like
// storeLeadF[li] = Id in Diode::Master::loadDAEVectors (N_DEV_Diode.C:2111-2118).
enum Kind { K_F = 0, K_Q, K_FLIM, K_QLIM, K_DFDX, K_DQDX, K_LEADF, K_LEADQ, NK };

inline bool kind_is_matrix(int k) { return k == K_DFDX || k == K_DQDX; }
inline bool kind_is_assign(int k) { return k == K_LEADF || k == K_LEADQ; }

// One stamp entry of an instance: local row node and local column node
// (vector kinds use `row` only; lead kinds use `row` as local store index).
struct Stamp {
  int16_t row, col;
};

// Parameter table row: name as written on a .MODEL / instance line, slot in
// the card vector, default value, and whether the value is a temperature
// given in Celsius (converted to kelvin on intake, DeviceEntity::setParam,
// N_DEV_DeviceEntity.C:285-300).
struct ParamDef {
  const char* name;
  int index;
  double def;
  int is_temp;
};

struct Card {
  int level = 1;
  std::vector<double> v;
  std::vector<char> given;
  bool has(int idx) const { return given[idx] != 0; }
};

// Fixed-order segmented scatter plan of one kind.  Target t receives
// vals[src[seg[t]]], vals[src[seg[t]+1]], ... added to out[target[t]] in that
// order, which is instance order then stamp order -- the order in which the
// reference's per-instance loops execute their `+=` (DeviceMaster default
// loops, N_DEV_DeviceMaster.h:643-725).
struct Plan {
  int E = 0;                       // value slots per instance
  int64_t n_targets = 0, n_src = 0;
  int32_t* d_target = nullptr;     // null when the targets are t0, t0+1, ... (contiguous)
  int32_t t0 = 0;
  bool idx32 = false;              // seg/src held as int32 (every index < 2^31)
  void* d_seg = nullptr;           // n_targets + 1 segment starts (int32 or int64)
  void* d_src = nullptr;           // n_src value-buffer slots (int32 or int64)
  int32_t* d_long = nullptr;       // targets whose segments go to the one-wave kernel
  int64_t n_long = 0;
  // rails: segments summed in fixed blocks, then the block sums in order
  int32_t* d_rail_t = nullptr;     // rail targets
  int32_t* d_rail_c = nullptr;     // rails + 1 offsets into the blocks
  void* d_chunk_j0 = nullptr;      // block source ranges (plan index type)
  void* d_chunk_j1 = nullptr;
  double* d_partial = nullptr;     // block sums
  int64_t n_rails = 0, n_chunks = 0;
  double* d_vals = nullptr;        // E * n, entry-major: slot e of instance i at e*n + i
  // Pattern plan: instance i feeds S targets at t0 + i * pP + offset (the
  // same offsets for every instance, the last at plast < pP), and entry e of
  // every instance goes to offset pslot[e] (-1: none).  The scatter then
  // needs no segment or source arrays.
  bool pattern = false;
  int pS = 0;
  int pP = 0;
  int plast = 0;
  int pblock = 0;                  // instances per block of k_scatter_pattern
  int32_t* d_pslot = nullptr;      // E offsets
  // The fused load needs every target's contributions to come from the
  // instances of one block (its block plan is the group's flat plan).
  bool block_private = true;
  int loff = 0;                    // first LDS row of this kind
};

// Flat block plan of the fused load, every kind's targets together.  For
// each block of fused_B instances, kind by kind (targets ascending), one
// record per target: {target | kind << kKindShift, first source}; the next
// record's first source ends it (a last record closes the list).  Sources are
// LDS slots (loff + e) * B + lane.  Targets of more than kWaveSegment terms
// (lead assignments aside) are listed again per block as
// {target | kind << kKindShift, first, end, 0}: a whole wavefront sums each.
constexpr int kKindShift = 29;
constexpr uint32_t kTargetMask = (1u << kKindShift) - 1;
constexpr int32_t kWaveSegment = 16;
struct FRec { int32_t tk, j0; };
struct __attribute__((aligned(16))) FLong { int32_t tk, j0, j1, pad; };

struct Group;

// Per-model back end.
struct ModelImpl {
  int type;
  const char* name;
  const ParamDef* mparams; int n_mparams;   // model-card parameters (+aliases)
  int n_mslots;                             // card vector length
  const ParamDef* iparams; int n_iparams;   // instance parameters (+aliases)
  int n_islots;
  int max_nodes;            // length of the full local node list
  int max_sto, max_sta;     // local store / state slots used by the kernels
  int max_entries[NK];      // value slots per instance per kind
  int (*setup)(Group&);     // processParams / updateTemperature on the host
  // map registered LIDs (ext then int) to the full local node list
  void (*resolve_nodes)(const Group&, int64_t inst, const int32_t* reg, int32_t* full);
  // stamp list of one instance for one kind (length <= max_entries[kind])
  void (*stamps)(const Group&, int64_t inst, int kind, std::vector<Stamp>& out);
  int (*launch_update)(Group&, const xa_solver_state&, const xa_vectors&, bool fused);
  int (*launch_secondary)(Group&, const xa_solver_state&, const xa_vectors&);
  bool (*skip_kind)(const Group&, const xa_solver_state&, int kind);  // may be null
  const char* const* state_names; int n_state;  // persistent per-instance doubles
  double (*traffic)(const Group&, int phase);   // algorithmic bytes per instance
  int fused_block;          // threads per block of the fused load (0: no fused load)
  bool fused_at_scale;      // fused load also when the launch fills the GPU
  bool fused_small;         // fused load when the launch is a few blocks
  // pattern-fused load (one kernel staging the outputs of pattern plans in
  // LDS); null when the model has none
  int (*launch_pattern)(Group&, const xa_solver_state&, const xa_vectors&);
  // per-instance columns the update reads from its card (null: none)
  int (*card_columns)(const Group&);
  // workgroups of the persistent update that fit the GPU (null: no
  // persistent update)
  int (*persistent_capacity)(const Group&);
};

const ModelImpl* model_diode();
const ModelImpl* model_resistor();
const ModelImpl* model_capacitor();
const ModelImpl* model_vsrc();
const ModelImpl* model_mosfet1();
const ModelImpl* model_bjt();
const ModelImpl* model_bsim4();

struct Ctx;

enum Phase { PH_CREATED = 0, PH_SETUP, PH_LIDS, PH_CSR, PH_FINAL };

struct Group {
  Ctx* ctx = nullptr;
  const ModelImpl* impl = nullptr;
  xa_device_options opt{};
  int phase = PH_CREATED;

  std::vector<Card> models;
  int64_t n = 0;
  std::vector<int32_t> inst_model;
  // instance parameters: card-like per instance, stored per slot column
  std::vector<std::vector<double>> ival;   // [slot][inst]
  std::vector<char> igiven;                // [slot] given for the whole group
  std::vector<int32_t> lead;               // [inst] lead current requested

  // filled by setup
  std::vector<int32_t> n_ext, n_int, n_sta, n_sto;
  int n_dfield = 0, n_ifield = 0;
  std::vector<double> dfield;    // [field][inst]
  std::vector<int32_t> ifield;   // [field][inst]
  std::vector<double> mcard;     // device model cards, model-major
  int mcard_stride = 0;
  // per-instance columns served from the card table instead (bit f: column f
  // holds one value per card; models that use it put the value in the card)
  uint64_t ucol[3] = {0, 0, 0};
  int n_state = 0;
  std::vector<double> state_init;  // [state][inst]
  // Value slots per instance of each kind (the stamp lists' length): the
  // model's max_entries, unless setup chose a longer list for this group.
  int ent[NK] = {0, 0, 0, 0, 0, 0, 0, 0};
  // The group loads only through the four calls (its kernels have no fused
  // or pattern-fused form), and a model-specific variant number setup chose.
  bool split_only = false;
  int variant = 0;
  // workgroups of the persistent stored-mode update (0: one per tile)
  int persist_blocks = 0;

  // filled by register_lids / csr
  std::vector<int32_t> nodes;    // [local node][inst]
  std::vector<int32_t> sto_lid;  // [local sto][inst]
  std::vector<int32_t> sta_lid;  // [local sta][inst]
  std::vector<int64_t> reg_off;  // prefix offsets of registered node lids
  std::vector<int32_t> tgt[NK];  // [inst*E + e] target, -1 none

  // device copies
  double* d_dfield = nullptr;
  int32_t* d_ifield = nullptr;
  double* d_mcard = nullptr;
  int32_t* d_nodes = nullptr;
  int32_t* d_sto = nullptr;
  int32_t* d_sta = nullptr;
  double* d_state = nullptr;
  int32_t* d_fail = nullptr;
  Plan plan[NK];
  int32_t* d_fblk = nullptr;   // flat block plan: blocks + 1 offsets into d_frec
  FRec* d_frec = nullptr;      // records (+ the closing one)
  uint16_t* d_fsrc = nullptr;  // LDS slots
  int32_t* d_flblk = nullptr;  // blocks + 1 offsets into d_flong
  FLong* d_flong = nullptr;    // long targets
  bool fused_ok = false;     // the fused load's block plans exist
  size_t fused_lds = 0;      // its LDS bytes per block
  int fused_rows = 0;        // LDS rows of the fused load (stamp rows of live kinds)
  int fused_B = 0;           // instances per block the block plans were built for
  int split_block = 256;     // threads per block of the stored-mode update kernel
  int route = 0;             // xa_load_all: 0 size/model policy, 1 four calls, 2 fused,
                             // 3 pattern-fused
  bool pattern_ok = false;   // every kind with targets has a pattern plan, no lead kinds
  int pattern_rows = 0;      // LDS doubles per instance of the pattern-fused load
  int pattern_B = 0;         // its instances per block
  // Per-block card tables of the stored-mode update: for every block of
  // card_block instances, the number of distinct models it uses and the
  // models (at most kBlockCards; kBlockCards + 1 ints per block), and each
  // instance's slot in its block's table.
  int card_block = 0;        // block size the tables were built for (0: none)
  bool card_tables = false;  // every block uses at most kBlockCards models
  int32_t* d_blk_cards = nullptr;
  uint8_t* d_card_slot = nullptr;

  // per-phase HIP event pairs recorded while timing is on
  bool timing = false;
  std::vector<hipEvent_t> ev[3];
  size_t n_ev[3] = {0, 0, 0};
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double* scratch = nullptr;   // this device's address of the no-alloc load's scratch
};

// error reporting
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
#define XA_HIP(call) do { hipError_t _e = (call); if (_e != hipSuccess) return ::xa::hip_check(_e, #call); } while (0)

// block size used by the per-instance kernels
constexpr int kBlock = 256;
// distinct model cards a block of the stored-mode update stages in LDS
constexpr int kBlockCards = 16;
// LDS of one workgroup on gfx950
constexpr size_t kMaxLds = 160 * 1024;
// A load of at most this many blocks is launch-latency bound: xa_load_all
// then takes the one-kernel fused form whatever the model prefers at scale.
constexpr int64_t kSmallLaunchBlocks = 512;
inline unsigned grid_for(int64_t n, int block = kBlock) {
  return (unsigned)((n + block - 1) / block);
}

}  // namespace xa
