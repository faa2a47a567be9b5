/* or_core.c -- oracle group plumbing (test infrastructure only): parameter
 * intake by name with given flags and Celsius->kelvin conversion of
 * temperature parameters (DeviceEntity::setParam, N_DEV_DeviceEntity.C:285-300),
 * LID registration and CSR slot lookup (CktGraphBasic::registerJacLIDswithDevs,
 * N_TOP_CktGraphBasic.C:1092-1136). */
#include <ctype.h>
#include <stdio.h>
#include "or_internal.h"

static __thread char or_err[512];
int or_threads = 1;
int or_mag = 0;
void or_set_magnitude_mode(int on) { or_mag = on != 0; }
int or_libm_ulps = 0;
uint64_t or_libm_seed = 0;
void or_set_libm_perturbation(int ulps, uint64_t seed) {
  or_libm_ulps = ulps > 0 ? ulps : 0;
  or_libm_seed = seed;
}

/* SplitMix64 finaliser over the argument bits and the seed */
double or_perturb(double r, double a, double b) {
  uint64_t ua, ub, h;
  int64_t k;
  int j;
  if (!isfinite(r) || r == 0.0) return r;
  memcpy(&ua, &a, 8);
  memcpy(&ub, &b, 8);
  h = ua ^ (ub * 0x9E3779B97F4A7C15ull) ^ or_libm_seed;
  h += 0x9E3779B97F4A7C15ull;
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  /* a nonzero offset: every call moves, so every sensitive call shows */
  k = (int64_t)(h % (uint64_t)(2 * or_libm_ulps)) - or_libm_ulps;
  if (k >= 0) ++k;
  for (j = 0; j < k; ++j) r = nextafter(r, INFINITY);
  for (j = 0; j > k; --j) r = nextafter(r, -INFINITY);
  return r;
}

/* Snapshot / restore of every instance record (the values the reference
 * keeps in its Instance objects between calls). */
int or_group_save_state(or_group g) {
  size_t bytes;
  if (!g || !g->inst || !g->ops->inst_size) return or_fail(XA_ERR_STATE, "oracle: nothing to save");
  bytes = g->ops->inst_size * (size_t)(g->n > 0 ? g->n : 1);
  if (!g->saved) g->saved = malloc(bytes);
  memcpy(g->saved, g->inst, bytes);
  return XA_OK;
}
int or_group_restore_state(or_group g) {
  if (!g || !g->saved) return or_fail(XA_ERR_STATE, "oracle: no saved state");
  memcpy(g->inst, g->saved, g->ops->inst_size * (size_t)(g->n > 0 ? g->n : 1));
  return XA_OK;
}

int or_fail(int code, const char* msg) {
  snprintf(or_err, sizeof(or_err), "%s", msg);
  return code;
}
const char* or_last_error(void) { return or_err; }
void or_set_threads(int n) { or_threads = n > 0 ? n : 1; }

/* The evaluation's math functions as this build resolves them (the C
 * library's, or in the parity build the portable ones): 0 exp(a), 1 log(a),
 * 2 pow(a, b), 3 tan(a), 4 sin(a).  For the accuracy tests of the portable functions. */
double or_math(int which, double a, double b) {
  switch (which) {
    case 0: return OR_EXP(a);
    case 1: return OR_LOG(a);
    case 2: return OR_POW(a, b);
    case 3: return OR_TAN(a);
    case 4: return OR_SIN(a);
  }
  return NAN;
}

/* DeviceOptions defaults (N_DEV_DeviceOptions.C:66-140) */
void or_default_options(xa_device_options* o) {
  memset(o, 0, sizeof(*o));
  o->gmin = 1.0e-12;
  o->tnom = OR_REFTEMP;
  o->temp = OR_REFTEMP;
  o->defl = 1.0e-4;
  o->defw = 1.0e-4;
  o->excessPhaseScalar1 = 1.0;
  o->excessPhaseScalar2 = 1.0;
  o->voltageLimiterFlag = 1;
}

static const or_model_ops* ops_for(int type) {
  switch (type) {
    case XA_DIODE: return &or_diode_ops;
    case XA_RESISTOR: return &or_resistor_ops;
    case XA_CAPACITOR: return &or_capacitor_ops;
    case XA_VSOURCE: return &or_vsrc_ops;
    case XA_MOSFET1: return &or_mosfet1_ops;
    case XA_BJT: return &or_bjt_ops;
    case XA_BSIM4: return &or_bsim4_ops;
  }
  return NULL;
}

static int name_eq(const char* a, const char* b) {
  for (; *a && *b; ++a, ++b)
    if (toupper((unsigned char)*a) != toupper((unsigned char)*b)) return 0;
  return *a == 0 && *b == 0;
}

int or_group_create(int model_type, const xa_device_options* opt, or_group* out) {
  const or_model_ops* ops = ops_for(model_type);
  struct or_group_s* g;
  if (!ops || !ops->setup) return or_fail(XA_ERR_ARG, "oracle: unknown model type");
  g = (struct or_group_s*)calloc(1, sizeof(*g));
  g->type = model_type;
  g->ops = ops;
  if (opt) g->opt = *opt;
  else or_default_options(&g->opt);
  g->ival = (double**)calloc(ops->n_islots, sizeof(double*));
  g->igiven = (char*)calloc(ops->n_islots, 1);
  *out = g;
  return XA_OK;
}

int or_group_destroy(or_group g) {
  int k;
  if (!g) return XA_OK;
  if (g->ops->destroy) g->ops->destroy(g);
  for (k = 0; k < g->nmodels; ++k) { free(g->models[k].v); free(g->models[k].given); }
  free(g->models);
  for (k = 0; k < g->ops->n_islots; ++k) free(g->ival[k]);
  free(g->ival); free(g->igiven); free(g->inst_model); free(g->lead);
  free(g->n_ext); free(g->n_int); free(g->n_sta); free(g->n_sto);
  free(g->inst); free(g->mdata); free(g->saved);
  free(g);
  return XA_OK;
}

int or_group_add_model(or_group g, int level, int n_params, const char* const* names,
                       const double* values, int* model_index) {
  const or_model_ops* o = g->ops;
  or_card c;
  int p, k;
  char* seen;
  c.level = level;
  c.v = (double*)calloc(o->n_mslots, sizeof(double));
  c.given = (char*)calloc(o->n_mslots, 1);
  seen = (char*)calloc(o->n_mslots, 1);
  for (k = 0; k < o->n_mparams; ++k)
    if (!seen[o->mparams[k].slot]) { c.v[o->mparams[k].slot] = o->mparams[k].def; seen[o->mparams[k].slot] = 1; }
  free(seen);
  for (p = 0; p < n_params; ++p) {
    const or_param* d = NULL;
    for (k = 0; k < o->n_mparams; ++k)
      if (name_eq(names[p], o->mparams[k].name)) { d = &o->mparams[k]; break; }
    if (!d) { free(c.v); free(c.given); return or_fail(XA_ERR_ARG, "oracle: unknown model parameter"); }
    c.v[d->slot] = values[p] + (d->is_temp ? OR_CtoK : 0.0);
    c.given[d->slot] = 1;
  }
  g->models = (or_card*)realloc(g->models, sizeof(or_card) * (g->nmodels + 1));
  g->models[g->nmodels] = c;
  if (model_index) *model_index = g->nmodels;
  g->nmodels++;
  return XA_OK;
}

int or_group_add_instances(or_group g, int n, const int* model_index, int n_params,
                           const char* const* names, const double* const* values) {
  const or_model_ops* o = g->ops;
  int64_t n0 = g->n, i;
  int p, k, s;
  const or_param** defs = (const or_param**)calloc(n_params > 0 ? n_params : 1, sizeof(*defs));
  for (p = 0; p < n_params; ++p) {
    for (k = 0; k < o->n_iparams; ++k)
      if (name_eq(names[p], o->iparams[k].name)) { defs[p] = &o->iparams[k]; break; }
    if (!defs[p]) { free(defs); return or_fail(XA_ERR_ARG, "oracle: unknown instance parameter"); }
  }
  g->n = n0 + n;
  g->inst_model = (int*)realloc(g->inst_model, sizeof(int) * g->n);
  g->lead = (int32_t*)realloc(g->lead, sizeof(int32_t) * g->n);
  for (i = 0; i < n; ++i) {
    g->inst_model[n0 + i] = model_index ? model_index[i] : 0;
    g->lead[n0 + i] = 0;
  }
  for (s = 0; s < o->n_islots; ++s) {
    double def = 0.0;
    for (k = 0; k < o->n_iparams; ++k)
      if (o->iparams[k].slot == s) { def = o->iparams[k].def; break; }
    g->ival[s] = (double*)realloc(g->ival[s], sizeof(double) * g->n);
    for (i = n0; i < g->n; ++i) g->ival[s][i] = def;
  }
  for (p = 0; p < n_params; ++p) {
    g->igiven[defs[p]->slot] = 1;
    for (i = 0; i < n; ++i)
      g->ival[defs[p]->slot][n0 + i] = values[p][i] + (defs[p]->is_temp ? OR_CtoK : 0.0);
  }
  free(defs);
  return XA_OK;
}

int or_group_request_lead_current(or_group g, const int32_t* flags) {
  int64_t i;
  for (i = 0; i < g->n; ++i) g->lead[i] = flags ? (flags[i] != 0) : 0;
  return XA_OK;
}

int or_group_setup(or_group g) {
  g->n_ext = (int*)calloc(g->n + 1, sizeof(int));
  g->n_int = (int*)calloc(g->n + 1, sizeof(int));
  g->n_sta = (int*)calloc(g->n + 1, sizeof(int));
  g->n_sto = (int*)calloc(g->n + 1, sizeof(int));
  g->phase = 1;
  return g->ops->setup(g);
}

int or_group_counts(or_group g, int* n_ext, int* n_int, int* n_sta, int* n_sto) {
  int64_t i;
  for (i = 0; i < g->n; ++i) {
    if (n_ext) n_ext[i] = g->n_ext[i];
    if (n_int) n_int[i] = g->n_int[i];
    if (n_sta) n_sta[i] = g->n_sta[i];
    if (n_sto) n_sto[i] = g->n_sto[i];
  }
  return XA_OK;
}

int or_group_register_lids(or_group g, const int32_t* node_lids, const int32_t* sta_lids,
                           const int32_t* sto_lids) {
  int64_t i, on = 0, osta = 0, osto = 0;
  for (i = 0; i < g->n; ++i) {
    g->ops->register_lids(g, i, node_lids + on, sta_lids ? sta_lids + osta : NULL,
                          sto_lids ? sto_lids + osto : NULL);
    on += g->n_ext[i] + g->n_int[i];
    osta += g->n_sta[i];
    osto += g->n_sto[i];
  }
  g->phase = 2;
  return XA_OK;
}

int or_group_jac_pairs(or_group g, int64_t* count, int32_t* rows, int32_t* cols) {
  int64_t i, c = 0;
  for (i = 0; i < g->n; ++i)
    c += g->ops->jac_pairs(g, i, rows ? rows + c : NULL, cols ? cols + c : NULL);
  if (count) *count = c;
  return XA_OK;
}

int64_t or_slot(const struct or_group_s* g, int32_t row, int32_t col) {
  int32_t k;
  if (row < 0 || col < 0) return -1;
  for (k = g->row_ptr[row]; k < g->row_ptr[row + 1]; ++k)
    if (g->col_idx[k] == col) return k;
  return -2;
}

int or_group_register_csr(or_group g, int n_rows, const int32_t* row_ptr,
                          const int32_t* col_idx) {
  int64_t i;
  int rc;
  g->n_rows = n_rows;
  g->row_ptr = row_ptr;
  g->col_idx = col_idx;
  for (i = 0; i < g->n; ++i)
    if ((rc = g->ops->register_jac(g, i)) != XA_OK) break;
  g->row_ptr = NULL;
  g->col_idx = NULL;
  g->phase = 3;
  return i == g->n ? XA_OK : or_fail(XA_ERR_ARG, "oracle: stamp missing from CSR structure");
}

int or_group_finalize(or_group g) { g->phase = 4; return XA_OK; }

int or_update_state(or_group g, const xa_solver_state* ss, const xa_vectors* v) {
  return g->ops->update_state(g, ss, v);
}
int or_update_secondary_state(or_group g, const xa_solver_state* ss, const xa_vectors* v) {
  return g->ops->update_secondary ? g->ops->update_secondary(g, ss, v) : XA_OK;
}
int or_load_vectors(or_group g, const xa_solver_state* ss, const xa_vectors* v) {
  return g->ops->load_vectors(g, ss, v);
}
int or_load_matrices(or_group g, const xa_solver_state* ss, const xa_vectors* v) {
  return g->ops->load_matrices(g, ss, v);
}
int or_read_intermediate(or_group g, const char* name, double* out) {
  return g->ops->read ? g->ops->read(g, name, out) : or_fail(XA_ERR_ARG, "no intermediates");
}
