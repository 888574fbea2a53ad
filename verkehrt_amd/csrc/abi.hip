// The C ABI of the hip variant (include/vk_hip.h): the kernel table by reference name, the
// reference's argument checks and error texts, reducer / scanner / builder states on the
// device, and the trace lines of src/exec.rs:390-420.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "vk_internal.h"

namespace vk {
Ctx *create_ctx(int device);
void destroy_ctx(Ctx *c);

namespace {

// ------------------------------------------------------------------ the table

enum InKind { IN_ARRAY, IN_SCALAR, IN_EITHER };
struct In {
  InKind k;
  int32_t t;
};

enum Impl {
  R_CANON, R_SUM_INT, R_PROD_INT, R_COUNT_CMP, R_COUNT_BOOL, R_EXTREME, R_ARG, R_BITS, R_FSUM_INT,
  S_COMPRESS, S_FILTER, S_WHERE,
  SC_SCAN, SC_FSCAN, SC_PREV,
  W_GATHER, W_ARGSORT, W_GROUP_ID, W_GROUP_ID2, W_HISTOGRAM, W_REDUCE_INTO, W_RAND_F64, W_RAND_I64,
  W_IOTA, W_LENGTH, W_FILL,
  E_BINARY, E_UNARY, E_SELECT, E_DIVC
};

struct Def {
  std::string name;
  int cls;
  std::vector<In> in;
  bool out_array;
  int32_t out_t;
  int impl;
  int p0 = 0, p1 = 0;  // family parameters (cmp, op, kind, max flag, ...)
  int32_t t0 = 0, t1 = 0;  // element types the family works on
};

std::vector<Def> g_defs;
std::unordered_map<std::string, int> g_by_name;

const char *kCmp[6] = {"eq", "ne", "lt", "le", "gt", "ge"};

void add(Def d) {
  g_by_name[d.name] = static_cast<int>(g_defs.size());
  g_defs.push_back(std::move(d));
}

std::string tn(int32_t t) { return type_name(t); }

void build_table() {
  if (!g_defs.empty()) return;
  const int32_t ints[] = {VK_I64, VK_U64, VK_I32, VK_U8};
  const int32_t nums[] = {VK_F64, VK_F32, VK_I64, VK_U64, VK_I32, VK_U8};
  const int32_t all[] = {VK_F64, VK_F32, VK_I64, VK_U64, VK_I32, VK_U8, VK_BOOL};
  auto A = [](int32_t t) { return In{IN_ARRAY, t}; };
  auto S = [](int32_t t) { return In{IN_SCALAR, t}; };
  auto E = [](int32_t t) { return In{IN_EITHER, t}; };
  auto sum_t = [](int32_t t) { return t == VK_I32 ? VK_I64 : t == VK_U8 ? VK_U64 : t; };

  // reductions (src/kernels/scalar/reduce.rs:680-726)
  add({"reduce_add_f64", VK_REDUCTION, {A(VK_F64)}, false, VK_F64, R_CANON, 0, 0, VK_F64});
  add({"reduce_mul_f64", VK_REDUCTION, {A(VK_F64)}, false, VK_F64, R_CANON, 4, 0, VK_F64});
  add({"reduce_add_f32", VK_REDUCTION, {A(VK_F32)}, false, VK_F32, R_CANON, 0, 0, VK_F32});
  add({"reduce_mul_f32", VK_REDUCTION, {A(VK_F32)}, false, VK_F32, R_CANON, 4, 0, VK_F32});
  for (int32_t t : ints) {
    add({"reduce_add_" + tn(t), VK_REDUCTION, {A(t)}, false, sum_t(t), R_SUM_INT, 0, 0, t});
    add({"reduce_mul_" + tn(t), VK_REDUCTION, {A(t)}, false, t, R_PROD_INT, 0, 0, t});
  }
  for (int32_t t : nums) {
    add({"reduce_min_" + tn(t), VK_REDUCTION, {A(t)}, false, t, R_EXTREME, 0, 0, t});
    add({"reduce_max_" + tn(t), VK_REDUCTION, {A(t)}, false, t, R_EXTREME, 1, 0, t});
    add({"argmin_" + tn(t), VK_REDUCTION, {A(t)}, false, VK_I64, R_ARG, 0, 0, t});
    add({"argmax_" + tn(t), VK_REDUCTION, {A(t)}, false, VK_I64, R_ARG, 1, 0, t});
  }
  const char *bitn[3] = {"and", "or", "xor"};
  for (int32_t t : ints)
    for (int b = 0; b < 3; ++b)
      add({std::string("reduce_") + bitn[b] + "_" + tn(t), VK_REDUCTION, {A(t)}, false, t, R_BITS, b, 0, t});
  for (int b = 0; b < 3; ++b)
    add({std::string("reduce_") + bitn[b] + "_bool", VK_REDUCTION, {A(VK_BOOL)}, false, VK_BOOL, R_BITS, b, 0, VK_BOOL});
  add({"count", VK_REDUCTION, {A(VK_BOOL)}, false, VK_I64, R_COUNT_BOOL});
  add({"any", VK_REDUCTION, {A(VK_BOOL)}, false, VK_BOOL, R_BITS, 1, 0, VK_BOOL});
  add({"all", VK_REDUCTION, {A(VK_BOOL)}, false, VK_BOOL, R_BITS, 0, 0, VK_BOOL});
  // fused kernels (src/kernels/scalar/fused.rs:248-379): canonical float sums for f64 and f32,
  // wrapping widened sums for i64, u64 and i32
  for (int32_t t : {VK_F64, VK_F32}) {
    add({"dot_" + tn(t), VK_REDUCTION, {A(t), A(t)}, false, t, R_CANON, 1, 0, t});
    add({"sumsq_" + tn(t), VK_REDUCTION, {A(t)}, false, t, R_CANON, 2, 0, t});
    for (int c = 0; c < 6; ++c)
      add({std::string("condsum_") + kCmp[c] + "_" + tn(t), VK_REDUCTION, {A(t), S(t), S(t)}, false, t, R_CANON,
           3, c, t});
  }
  for (int32_t t : {VK_I64, VK_U64, VK_I32}) {
    add({"dot_" + tn(t), VK_REDUCTION, {A(t), A(t)}, false, sum_t(t), R_FSUM_INT, 1, 0, t});
    add({"sumsq_" + tn(t), VK_REDUCTION, {A(t)}, false, sum_t(t), R_FSUM_INT, 2, 0, t});
    for (int c = 0; c < 6; ++c)
      add({std::string("condsum_") + kCmp[c] + "_" + tn(t), VK_REDUCTION, {A(t), S(t), S(t)}, false, sum_t(t),
           R_FSUM_INT, 3, c, t});
  }
  for (int c = 0; c < 6; ++c) {
    for (int32_t t : {VK_F64, VK_F32, VK_I64, VK_U64, VK_I32}) {
      add({std::string("count_") + kCmp[c] + "_" + tn(t), VK_REDUCTION, {A(t), S(t)}, false, VK_I64,
           R_COUNT_CMP, c, 0, t});
      add({std::string("filter_") + kCmp[c] + "_" + tn(t), VK_SELECTION, {A(t), S(t)}, true, t,
           S_FILTER, c, 0, t});
    }
  }
  // selections (select.rs:229-257)
  for (int32_t t : all)
    add({"compress_" + tn(t), VK_SELECTION, {A(t), A(VK_BOOL)}, true, t, S_COMPRESS, 0, 0, t});
  add({"where", VK_SELECTION, {A(VK_BOOL)}, true, VK_I64, S_WHERE});
  // scans (scan.rs:386-439)
  for (int32_t t : ints) {
    add({"scan_add_" + tn(t), VK_SCAN, {A(t)}, true, sum_t(t), SC_SCAN, 0, 0, t});
    add({"scan_by_add_" + tn(t), VK_SCAN, {A(t), A(VK_BOOL)}, true, sum_t(t), SC_SCAN, 0, 1, t});
  }
  for (int32_t t : {VK_F64, VK_I64, VK_U64}) {
    add({"scan_min_" + tn(t), VK_SCAN, {A(t)}, true, t, SC_SCAN, 1, 0, t});
    add({"scan_max_" + tn(t), VK_SCAN, {A(t)}, true, t, SC_SCAN, 2, 0, t});
    add({"scan_by_min_" + tn(t), VK_SCAN, {A(t), A(VK_BOOL)}, true, t, SC_SCAN, 1, 1, t});
    add({"scan_by_max_" + tn(t), VK_SCAN, {A(t), A(VK_BOOL)}, true, t, SC_SCAN, 2, 1, t});
  }
  // Not a reference kernel (builtins.dd:50-54): the order defined by this backend.
  add({"scan_add_f64", VK_SCAN, {A(VK_F64)}, true, VK_F64, SC_FSCAN});
  for (int32_t t : all)  // prev_<t> (scan.rs:57-75)
    add({"prev_" + tn(t), VK_SCAN, {A(t), S(t)}, true, t, SC_PREV, 0, 0, t});
  // whole kernels (index.rs:277-326, group.rs:330-365, construct.rs:200-272)
  for (int32_t t : all)
    add({"gather_" + tn(t), VK_WHOLE, {A(t), A(VK_I64)}, true, t, W_GATHER, 0, 0, t});
  for (int32_t t : {VK_I64, VK_U64, VK_I32, VK_U8, VK_F64})
    add({"argsort_" + tn(t), VK_WHOLE, {A(t)}, true, VK_I64, W_ARGSORT, 0, 0, t});
  for (int32_t a : ints) {
    add({"group_id_" + tn(a), VK_WHOLE, {A(a)}, true, VK_I64, W_GROUP_ID, 0, 0, a});
    for (int32_t b : ints)
      add({"group_id2_" + tn(a) + "_" + tn(b), VK_WHOLE, {A(a), A(b)}, true, VK_I64, W_GROUP_ID2, 0, 0, a, b});
  }
  add({"histogram", VK_WHOLE, {A(VK_I64), S(VK_I64)}, true, VK_I64, W_HISTOGRAM});
  add({"reduce_into_add_i64", VK_WHOLE, {A(VK_I64), A(VK_I64), S(VK_I64)}, true, VK_I64, W_REDUCE_INTO, 1, 0, VK_I64});
  // Not a reference kernel (builtins.dd:67-71): the order defined by this backend.
  add({"reduce_into_add_f64", VK_WHOLE, {A(VK_I64), A(VK_F64), S(VK_I64)}, true, VK_F64, W_REDUCE_INTO, 4, 0, VK_F64});
  for (int32_t t : {VK_F64, VK_I64}) {
    add({"reduce_into_min_" + tn(t), VK_WHOLE, {A(VK_I64), A(t), S(VK_I64)}, true, t, W_REDUCE_INTO, 2, 0, t});
    add({"reduce_into_max_" + tn(t), VK_WHOLE, {A(VK_I64), A(t), S(VK_I64)}, true, t, W_REDUCE_INTO, 3, 0, t});
  }
  add({"rand_f64", VK_WHOLE, {S(VK_I64), S(VK_I64)}, true, VK_F64, W_RAND_F64});
  add({"rand_i64", VK_WHOLE, {S(VK_I64), S(VK_I64)}, true, VK_I64, W_RAND_I64});
  add({"iota", VK_WHOLE, {S(VK_I64)}, true, VK_I64, W_IOTA});
  for (int32_t t : all) {
    add({"length_" + tn(t), VK_WHOLE, {A(t)}, false, VK_I64, W_LENGTH, 0, 0, t});
    add({"fill_" + tn(t), VK_WHOLE, {S(VK_I64), S(t)}, true, t, W_FILL, 0, 0, t});
  }
  // elementwise (elementwise.rs:740-912) over this backend's element types
  const char *arith[5] = {"add", "sub", "mul", "min", "max"};
  const int arith_op[5] = {EW_ADD, EW_SUB, EW_MUL, EW_MIN, EW_MAX};
  for (int32_t t : nums) {
    for (int k = 0; k < 5; ++k)
      add({std::string(arith[k]) + "_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, t, E_BINARY, arith_op[k], 0, t});
    for (int c = 2; c < 6; ++c)  // lt le gt ge (register_num)
      add({std::string(kCmp[c]) + "_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, VK_BOOL, E_BINARY, EW_EQ + c, 0, t});
  }
  for (int32_t t : all) {  // eq, ne and select exist for every type (register_any)
    for (int c = 0; c < 2; ++c)
      add({std::string(kCmp[c]) + "_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, VK_BOOL, E_BINARY, EW_EQ + c, 0, t});
    add({"select_" + tn(t), VK_ELEMENTWISE, {E(VK_BOOL), E(t), E(t)}, true, t, E_SELECT, EW_SELECT, 0, t});
  }
  for (int32_t t : {VK_I64, VK_I32, VK_F64, VK_F32}) {  // register_signed
    add({"neg_" + tn(t), VK_ELEMENTWISE, {E(t)}, true, t, E_UNARY, EW_NEG, 0, t});
    add({"abs_" + tn(t), VK_ELEMENTWISE, {E(t)}, true, t, E_UNARY, EW_ABS, 0, t});
  }
  for (int32_t t : {VK_F64, VK_F32}) {  // register_float
    add({"div_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, t, E_BINARY, EW_DIV, 0, t});
    add({"sqrt_" + tn(t), VK_ELEMENTWISE, {E(t)}, true, t, E_UNARY, EW_SQRT, 0, t});
  }
  for (int32_t t : ints)  // register_div_const: the divisor is a scalar
    add({"div_const_" + tn(t), VK_ELEMENTWISE, {E(t), S(t)}, true, t, E_DIVC, EW_DIVC, 0, t});
  for (int32_t t : {VK_I32, VK_I64, VK_U8, VK_U64}) {  // register_bits
    add({"and_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, t, E_BINARY, EW_BAND, 0, t});
    add({"or_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, t, E_BINARY, EW_BOR, 0, t});
    add({"xor_" + tn(t), VK_ELEMENTWISE, {E(t), E(t)}, true, t, E_BINARY, EW_BXOR, 0, t});
    add({"not_" + tn(t), VK_ELEMENTWISE, {E(t)}, true, t, E_UNARY, EW_BNOT, 0, t});
  }
  add({"and_bool", VK_ELEMENTWISE, {E(VK_BOOL), E(VK_BOOL)}, true, VK_BOOL, E_BINARY, EW_AND, 0, VK_BOOL});
  add({"or_bool", VK_ELEMENTWISE, {E(VK_BOOL), E(VK_BOOL)}, true, VK_BOOL, E_BINARY, EW_OR, 0, VK_BOOL});
  add({"xor_bool", VK_ELEMENTWISE, {E(VK_BOOL), E(VK_BOOL)}, true, VK_BOOL, E_BINARY, EW_XOR, 0, VK_BOOL});
  add({"not_bool", VK_ELEMENTWISE, {E(VK_BOOL)}, true, VK_BOOL, E_UNARY, EW_NOT, 0, VK_BOOL});
  for (int32_t f : all)  // register_casts: every source to every numeric type
    for (int32_t t : nums)
      add({"cast_" + tn(f) + "_" + tn(t), VK_ELEMENTWISE, {E(f)}, true, t, E_UNARY, EW_CAST, 0, f});
}

const Def &def_of(const char *name) {
  build_table();
  auto it = g_by_name.find(name ? name : "");
  if (it == g_by_name.end()) usage_error(std::string("no hip kernel named ") + (name ? name : "(null)"));
  return g_defs[it->second];
}

// ------------------------------------------------------------ argument checks

std::string arg_ty(const vk_arg &a) {
  if (a.kind == VK_ROWS) return a.rows ? tn(a.rows->type) + "[]" : std::string("?[]");
  return tn(a.splat.type);
}

// arity_error, elementwise.rs:215-217
[[noreturn]] void arity(size_t want, size_t got) {
  usage_error("expected " + std::to_string(want) + " arguments, got " + std::to_string(got));
}

// kind_error, dispatch.rs:323-330
[[noreturn]] void kind_error(const In &s, const vk_arg &a) {
  const std::string t = tn(s.t), got = arg_ty(a);
  if (s.k == IN_ARRAY) usage_error("expected " + t + "[], got " + got);
  if (s.k == IN_SCALAR) usage_error("expected " + t + ", got " + got);
  usage_error("expected " + t + " or " + t + "[], got " + got);
}

bool accepts(const In &s, const vk_arg &a) {
  if (a.kind == VK_ROWS) return a.rows && (s.k == IN_ARRAY || s.k == IN_EITHER) && a.rows->type == s.t;
  return (s.k == IN_SCALAR || s.k == IN_EITHER) && a.splat.type == s.t;
}

// prepare / check_tile_args (dispatch.rs:268-321): arity, kinds, one row count.
// Returns the common row count, or -1 when every argument is a scalar.
int64_t check_args(const Def &d, const vk_arg *args, size_t nargs) {
  if (nargs != d.in.size()) arity(d.in.size(), nargs);
  int64_t rows = -1;
  for (size_t i = 0; i < nargs; ++i) {
    if (!accepts(d.in[i], args[i])) kind_error(d.in[i], args[i]);
    if (args[i].kind == VK_ROWS) {
      const int64_t n = static_cast<int64_t>(args[i].rows->len);
      if (rows >= 0 && rows != n)
        usage_error("length mismatch: " + std::to_string(rows) + " vs " + std::to_string(n));
      rows = n;
    }
  }
  return rows;
}

// rows_arg / splat_arg (elementwise.rs:178-213), for entries that check their own arguments.
const vk_array *rows_arg(const vk_arg &a, int32_t ty) {
  if (a.kind == VK_ROWS && a.rows && a.rows->type == ty) return a.rows;
  if (a.kind == VK_ROWS) usage_error("expected " + tn(ty) + "[], got " + arg_ty(a));
  usage_error("expected " + tn(ty) + "[], got " + tn(a.splat.type));
}
uint64_t splat_arg(const vk_arg &a, int32_t ty) {
  if (a.kind == VK_SPLAT && a.splat.type == ty) return a.splat.bits;
  usage_error("expected " + tn(ty) + ", got " + arg_ty(a));
}
// i64_arg (construct.rs:23-28) for whole kernels
int64_t i64_arg(const vk_arg &a) {
  if (a.kind == VK_SPLAT && a.splat.type == VK_I64) return static_cast<int64_t>(a.splat.bits);
  usage_error("expected i64, got " + arg_ty(a));
}
const vk_array *array_arg(const vk_arg &a, int32_t ty) {
  if (a.kind == VK_ROWS && a.rows && a.rows->type == ty) return a.rows;
  usage_error("expected " + tn(ty) + "[], got " + arg_ty(a));
}

vk_scalar scalar(int32_t t, uint64_t bits) { return vk_scalar{t, bits}; }

uint64_t dbits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

// Allocation checks of Array::try_build (value.rs:283-307) for `len` rows of `t`.
vk_array *alloc_result(Ctx &c, int32_t t, uint64_t len) { return new_array(c, t, len); }

// ---------------------------------------------------------------- trace

struct TraceScope {
  Ctx &c;
  const char *name;
  uint64_t n = 0, read = 0, write = 0;
  bool on;
  TraceScope(Ctx &cc, const char *nm, const vk_arg *args, size_t nargs) : c(cc), name(nm), on(cc.trace) {
    if (!on) return;
    for (size_t i = 0; i < nargs; ++i)
      if (args[i].kind == VK_ROWS && args[i].rows) {
        if (args[i].rows->len > n) n = args[i].rows->len;
        read += abytes(args[i].rows);
      }
    (void)hipEventRecord(c.ev0, c.stream);
  }
  void done(uint64_t write_bytes) {
    if (!on) return;
    write = write_bytes;
    (void)hipEventRecord(c.ev1, c.stream);
    (void)hipEventSynchronize(c.ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
    c.last_us = ms * 1000.0;
    const double secs = ms / 1000.0;
    const double gbps = secs > 0 ? double(read + write) / secs / 1e9 : 0.0;
    char line[512];
    std::snprintf(line, sizeof line, "trace %s isa=hip n=%" PRIu64 " read=%" PRIu64 " write=%" PRIu64
                  " time=%.1f gbps=%.2f\n", name, n, read, write, c.last_us, gbps);
    c.trace_buf += line;
  }
};

}  // namespace
}  // namespace vk

// ------------------------------------------------------------ device states

struct vk_reducer {
  vk::Ctx *c;
  int def;
  void *state;  // device: canon acc[16] + tail[16]; others up to 3 words
  uint64_t consumed;
  int tail_len;
};

struct vk_scanner {
  vk::Ctx *c;
  int def;
  uint64_t *carry;  // device {acc bits, has}; f64 scan: 16 accumulators
  uint64_t consumed;
};

struct vk_builder {
  vk::Ctx *c;
  int32_t type;
  void *ptr;
  uint64_t cap, len;
};

namespace vk {
namespace {

ReduceSpec reduce_spec(const Def &d, uint64_t base) {
  ReduceSpec s{};
  s.type = d.t0;
  s.base_row = base;
  switch (d.impl) {
    case R_SUM_INT: s.kind = ReduceSpec::SUM_INT; break;
    case R_PROD_INT: s.kind = ReduceSpec::PROD_INT; break;
    case R_COUNT_CMP: s.kind = ReduceSpec::COUNT_CMP; s.cmp = d.p0; break;
    case R_COUNT_BOOL: s.kind = ReduceSpec::COUNT_BOOL; s.type = VK_U8; break;
    case R_EXTREME: s.kind = ReduceSpec::EXTREME; s.want_max = d.p0 != 0; break;
    case R_ARG: s.kind = ReduceSpec::ARG_EXTREME; s.want_max = d.p0 != 0; break;
    case R_FSUM_INT: s.kind = ReduceSpec::FSUM_INT; s.fkind = d.p0; s.cmp = d.p1; break;
    default: s.kind = ReduceSpec::BITS; s.bitop = d.p0; break;
  }
  return s;
}

vk_reducer *reducer_new(Ctx &c, int def) {
  auto *r = new vk_reducer{&c, def, nullptr, 0, 0};
  ctx_retain(c);
  r->state = tmp_alloc(c, 256);
  const Def &d = g_defs[def];
  if (d.impl == R_CANON) canon_init(c, d.p0 == 4, static_cast<double *>(r->state));
  else reduce_state_init(c, reduce_spec(d, 0), r->state);
  return r;
}

// Reducer::update with the entry's own argument checks (fused.rs, reduce.rs Fold::update).
void reducer_update(vk_reducer *r, const vk_arg *args, size_t nargs) {
  Ctx &c = *r->c;
  const Def &d = g_defs[r->def];
  if (d.impl == R_CANON) {
    const int kind = d.p0;
    const int32_t ft = d.t0;  // VK_F64 or VK_F32
    const vk_array *x = nullptr, *y = nullptr;
    CanonSpec cs{kind, d.p1, 0, 0};
    auto splat_val = [&](const vk_arg &a) {
      const uint64_t b = splat_arg(a, ft);
      if (ft == VK_F32) {
        float f;
        const uint32_t w = static_cast<uint32_t>(b);
        std::memcpy(&f, &w, 4);
        return static_cast<double>(f);
      }
      double v;
      std::memcpy(&v, &b, 8);
      return v;
    };
    if (kind == 1) {
      if (nargs != 2) arity(2, nargs);
      x = rows_arg(args[0], ft);
      y = rows_arg(args[1], ft);
      if (x->len != y->len)
        usage_error("length mismatch: " + std::to_string(x->len) + " vs " + std::to_string(y->len));
    } else if (kind == 3) {
      if (nargs != 3) arity(3, nargs);
      x = rows_arg(args[0], ft);
      cs.c = splat_val(args[1]);
      cs.e = splat_val(args[2]);
    } else {
      if (nargs != 1) arity(1, nargs);
      x = rows_arg(args[0], ft);
    }
    // CanonicalAcc::feed_blocks (reduce.rs:284-308)
    if (r->tail_len != 0) runtime_error("canonical update after a partial block");
    const uint64_t n = x->len, blocks = n / 16, rest = n % 16;
    double *acc = static_cast<double *>(r->state);
    if (ft == VK_F32) {
      const float *xp = static_cast<const float *>(dptr(x));
      const float *yp = y ? static_cast<const float *>(dptr(y)) : nullptr;
      canon32_blocks(c, cs, xp, yp, blocks, acc);
      canon32_rest(c, cs, xp + blocks * 16, yp ? yp + blocks * 16 : nullptr, rest, acc + 16);
    } else {
      const double *xp = static_cast<const double *>(dptr(x));
      const double *yp = y ? static_cast<const double *>(dptr(y)) : nullptr;
      canon_blocks(c, cs, xp, yp, blocks, acc);
      canon_rest(c, cs, xp + blocks * 16, yp ? yp + blocks * 16 : nullptr, rest, acc + 16);
    }
    r->tail_len = static_cast<int>(rest);
    r->consumed += n;
    return;
  }
  const vk_array *x;
  if (d.impl == R_COUNT_CMP) {
    if (nargs != 2) arity(2, nargs);
    x = rows_arg(args[0], d.t0);
    ReduceSpec s = reduce_spec(d, r->consumed);
    s.cbits = splat_arg(args[1], d.t0);
    reduce_rows(c, s, dptr(x), x->len, r->state);
  } else if (d.impl == R_FSUM_INT) {
    ReduceSpec s = reduce_spec(d, r->consumed);
    if (d.p0 == 1) {  // Dot::update (fused.rs:77-98)
      if (nargs != 2) arity(2, nargs);
      x = rows_arg(args[0], d.t0);
      const vk_array *y = rows_arg(args[1], d.t0);
      if (x->len != y->len)
        usage_error("length mismatch: " + std::to_string(x->len) + " vs " + std::to_string(y->len));
      reduce_rows2(c, s, dptr(x), dptr(y), x->len, r->state);
    } else if (d.p0 == 3) {  // CondSum::update (fused.rs:186-203)
      if (nargs != 3) arity(3, nargs);
      x = rows_arg(args[0], d.t0);
      s.cbits = splat_arg(args[1], d.t0);
      s.ebits = splat_arg(args[2], d.t0);
      reduce_rows(c, s, dptr(x), x->len, r->state);
    } else {  // SumSq::update (fused.rs:113-131)
      if (nargs != 1) arity(1, nargs);
      x = rows_arg(args[0], d.t0);
      reduce_rows(c, s, dptr(x), x->len, r->state);
    }
  } else {
    if (nargs != 1) arity(1, nargs);
    x = rows_arg(args[0], d.in[0].t);
    reduce_rows(c, reduce_spec(d, r->consumed), dptr(x), x->len, r->state);
  }
  r->consumed += x->len;
}

// The scalar of a finished non-canonical reduction state (Reducer::finish).
vk_scalar finish_words(const Def &d, const uint64_t *w) {
  switch (d.impl) {
    case R_EXTREME:
      if (!w[1]) runtime_error(std::string("reduce_") + (d.p0 ? "max" : "min") + " of an empty array");
      return scalar(d.out_t, w[0]);
    case R_ARG: {
      const int64_t nan = static_cast<int64_t>(w[2]), idx = static_cast<int64_t>(w[1]);
      if (nan != INT64_MAX) return scalar(VK_I64, static_cast<uint64_t>(nan));
      if (idx == INT64_MAX) runtime_error(std::string(d.p0 ? "argmax" : "argmin") + " of an empty array");
      return scalar(VK_I64, static_cast<uint64_t>(idx));
    }
    case R_PROD_INT:
    case R_BITS: {
      uint64_t v = w[0];
      if (d.out_t == VK_BOOL) v = v != 0;
      else if (type_width(d.out_t) < 8) v &= (uint64_t(1) << (8 * type_width(d.out_t))) - 1;
      return scalar(d.out_t, v);
    }
    default: return scalar(d.out_t, w[0]);
  }
}

vk_scalar reducer_finish(vk_reducer *r) {
  Ctx &c = *r->c;
  const Def &d = g_defs[r->def];
  if (d.impl == R_CANON) {
    double *acc = static_cast<double *>(r->state);
    if (d.t0 == VK_F32)
      canon32_finish(c, d.p0 == 4, acc, acc + 16, r->tail_len, reinterpret_cast<uint64_t *>(c.scratch.slots));
    else
      canon_finish(c, d.p0 == 4, acc, acc + 16, r->tail_len, reinterpret_cast<double *>(c.scratch.slots));
    uint64_t b;
    read_back(c, c.scratch.slots, 8, &b);
    return scalar(d.t0, b);
  }
  uint64_t w[3];
  read_back(c, r->state, 24, w);
  return finish_words(d, w);
}

void reducer_free(vk_reducer *r) {
  if (!r) return;
  Ctx *c = r->c;
  tmp_free(*c, r->state);
  delete r;
  ctx_release(c);
}

// Scanner::update (scan.rs:251-326) with tile_rows' checks (scan.rs:226-248).
void scanner_update(vk_scanner *s, const vk_arg *args, size_t nargs, vk_array *out) {
  Ctx &c = *s->c;
  const Def &d = g_defs[s->def];
  if (d.impl == SC_PREV) {
    // Prev::update (scan.rs:24-47)
    if (nargs != 2) arity(2, nargs);
    const vk_array *x = rows_arg(args[0], d.t0);
    const uint64_t fill = splat_arg(args[1], d.t0);
    if (!out || out->type != d.out_t) usage_error("bad destination: the output rows are not " + tn(d.out_t));
    if (out->len != x->len)
      usage_error("length mismatch: " + std::to_string(x->len) + " rows for a destination of " +
                  std::to_string(out->len) + " rows");
    prev_rows(c, type_width(d.t0), dptr(x), x->len, dptr(out), s->carry, fill);
    s->consumed += x->len;
    return;
  }
  const bool by = d.impl == SC_SCAN && d.p1;
  if (nargs != (by ? 2u : 1u)) arity(by ? 2 : 1, nargs);
  const vk_array *x = rows_arg(args[0], d.in[0].t);
  const vk_array *st = by ? rows_arg(args[1], VK_BOOL) : nullptr;
  if (!out || out->type != d.out_t) usage_error("bad destination: the output rows are not " + tn(d.out_t));
  if (out->len != x->len || (st && st->len != x->len))
    usage_error("length mismatch: " + std::to_string(x->len) + " rows for a destination of " +
                std::to_string(out->len) + " rows");
  if (x->len == 0) return;
  if (d.impl == SC_FSCAN) {
    if (s->consumed % 16 != 0)
      usage_error("scan_add_f64 takes tiles of multiples of 16 rows before the last");
    scan_add_f64(c, static_cast<const double *>(dptr(x)), x->len, static_cast<double *>(dptr(out)),
                 reinterpret_cast<double *>(s->carry));
  } else {
    ScanSpec sp{d.p0, d.t0, by};
    scan_rows(c, sp, dptr(x), st ? static_cast<const uint8_t *>(dptr(st)) : nullptr, x->len, dptr(out),
              s->carry);
  }
  s->consumed += x->len;
}

vk_scanner *scanner_new(Ctx &c, int def) {
  auto *s = new vk_scanner{&c, def, nullptr, 0};
  ctx_retain(c);
  s->carry = static_cast<uint64_t *>(tmp_alloc(c, 256));
  fill_u64(c, s->carry, 32, 0);
  return s;
}

void scanner_free(vk_scanner *s) {
  if (!s) return;
  Ctx *c = s->c;
  tmp_free(*c, s->carry);
  delete s;
  ctx_release(c);
}

// ------------------------------------------------------------------ builder

void builder_reserve(vk_builder *b, uint64_t extra) {
  const uint64_t need = b->len + extra;
  if (need <= b->cap) return;
  uint64_t cap = b->cap * 2 > need ? b->cap * 2 : need;
  const uint64_t w = type_width(b->type);
  void *p = b->c->pool->alloc(cap * w);
  if (!p) runtime_error("cannot allocate an array of " + std::to_string(cap) + " elements");
  if (b->ptr) {
    copy_d2d(*b->c, p, b->ptr, b->len * w);
    b->c->pool->free(b->ptr);
  }
  b->ptr = p;
  b->cap = cap;
}

// Appends the survivors of one tile (SelectionFn, kernels.rs:87-96).
void select_tile(Ctx &c, const Def &d, const vk_arg *args, size_t nargs, uint64_t base, vk_builder *b) {
  const int64_t rows = check_args(d, args, nargs);
  if (b->type != d.out_t)
    usage_error(d.name + " produces " + tn(d.out_t) + " rows, got a " + tn(b->type) + " builder");
  const uint64_t n = rows < 0 ? 0 : static_cast<uint64_t>(rows);
  if (n == 0) return;
  builder_reserve(b, n);
  void *out = static_cast<char *>(b->ptr) + b->len * type_width(b->type);
  uint64_t *d_cnt = c.scratch.slots + 16;
  SelectSpec s{};
  if (d.impl == S_COMPRESS) {
    s.kind = 0;
    s.type = d.t0;
    select_rows(c, s, dptr(args[0].rows), static_cast<const uint8_t *>(dptr(args[1].rows)), n, out, d_cnt);
  } else if (d.impl == S_FILTER) {
    s.kind = 1;
    s.type = d.t0;
    s.cmp = d.p0;
    s.cbits = args[1].splat.bits;
    select_rows(c, s, dptr(args[0].rows), nullptr, n, out, d_cnt);
  } else {
    s.kind = 2;
    s.type = VK_I64;
    s.base = base;
    select_rows(c, s, nullptr, static_cast<const uint8_t *>(dptr(args[0].rows)), n, out, d_cnt);
  }
  uint64_t got = 0;
  read_back(c, d_cnt, 8, &got);
  b->len += got;
}

vk_array *builder_take(vk_builder *b) {
  Ctx &c = *b->c;
  const uint64_t w = type_width(b->type);
  // A survivor count far below the reservation gets an exact buffer, so a 1% filter does
  // not hold the input's size in HBM.
  if (b->ptr && b->cap - b->len >= (uint64_t(16) << 20) / w && b->len <= b->cap / 4) {
    void *p = c.pool->alloc(b->len * w);
    if (p) {
      copy_d2d(c, p, b->ptr, b->len * w);
      c.pool->free(b->ptr);
      b->ptr = p;
      b->cap = b->len;
    }
  }
  if (!b->ptr) {
    b->ptr = c.pool->alloc(256);
    if (!b->ptr) runtime_error("cannot allocate an array of 0 elements");
  }
  auto *buf = new Buffer{&c, b->ptr, b->cap * w, true, 1};
  ctx_retain(c);
  auto *a = new vk_array{1, buf, 0, b->len, b->type};
  b->ptr = nullptr;
  b->cap = b->len = 0;
  return a;
}

// ------------------------------------------------------------ whole kernels

uint64_t count_arg(const vk_arg &a) {
  const int64_t n = i64_arg(a);
  if (n < 0) runtime_error("negative length " + std::to_string(n));
  return static_cast<uint64_t>(n);
}

// checked_keys (index.rs:113-126): bucket count, then the first key out of range.
uint64_t checked_keys(Ctx &c, const vk_array *keys, const vk_arg &nb) {
  const int64_t buckets = i64_arg(nb);
  if (buckets < 0) runtime_error("negative bucket count " + std::to_string(buckets));
  return static_cast<uint64_t>(buckets);
}

void bucket_error(Ctx &c, const vk_array *keys, uint64_t bad_row, int64_t buckets) {
  int64_t k;
  read_back(c, static_cast<const int64_t *>(dptr(keys)) + bad_row, 8, &k);
  runtime_error("bucket key " + std::to_string(k) + " out of range for " + std::to_string(buckets) +
                " buckets");
}

vk_value run_whole(Ctx &c, const Def &d, const vk_arg *args, size_t nargs) {
  vk_value v{VK_VALUE_ARRAY, nullptr, {0, 0}};
  switch (d.impl) {
    case W_GATHER: {
      if (nargs != 2) arity(2, nargs);
      const vk_array *xs = array_arg(args[0], d.t0);
      const vk_array *idx = array_arg(args[1], VK_I64);
      vk_array *out = alloc_result(c, d.t0, idx->len);
      uint64_t *d_bad = c.scratch.slots + 17;
      gather_rows(c, type_width(d.t0), dptr(xs), xs->len, static_cast<const int64_t *>(dptr(idx)),
                  idx->len, dptr(out), d_bad);
      uint64_t bad;
      read_back(c, d_bad, 8, &bad);
      if (bad != ~uint64_t(0)) {
        release(out);
        int64_t i;
        read_back(c, static_cast<const int64_t *>(dptr(idx)) + bad, 8, &i);
        runtime_error("gather index " + std::to_string(i) + " out of range for length " +
                      std::to_string(xs->len));
      }
      v.array = out;
      return v;
    }
    case W_ARGSORT: {
      if (nargs != 1) arity(1, nargs);
      const vk_array *xs = array_arg(args[0], d.t0);
      vk_array *out = alloc_result(c, VK_I64, xs->len);
      argsort(c, d.t0, dptr(xs), xs->len, static_cast<int64_t *>(dptr(out)));
      v.array = out;
      return v;
    }
    case W_GROUP_ID: case W_GROUP_ID2: {
      const bool pair = d.impl == W_GROUP_ID2;
      if (nargs != (pair ? 2u : 1u)) arity(pair ? 2 : 1, nargs);
      const vk_array *a = array_arg(args[0], d.t0);
      const vk_array *b = pair ? array_arg(args[1], d.t1) : nullptr;
      if (b && a->len != b->len)
        runtime_error("length mismatch: " + std::to_string(a->len) + " vs " + std::to_string(b->len));
      vk_array *out = alloc_result(c, VK_I64, a->len);
      group_id(c, d.t0, dptr(a), pair ? d.t1 : 0, b ? dptr(b) : nullptr, a->len,
               static_cast<int64_t *>(dptr(out)));
      v.array = out;
      return v;
    }
    case W_HISTOGRAM: case W_REDUCE_INTO: {
      const bool hist = d.impl == W_HISTOGRAM;
      if (nargs != (hist ? 2u : 3u)) arity(hist ? 2 : 3, nargs);
      const vk_array *vals = hist ? nullptr : array_arg(args[1], d.t0);  // vals first, index.rs:199
      const vk_array *keys = array_arg(args[0], VK_I64);
      const int64_t buckets = static_cast<int64_t>(checked_keys(c, keys, args[hist ? 1 : 2]));
      uint64_t *d_bad = c.scratch.slots + 18;
      uint64_t bad;
      const int op = hist ? 0 : d.p0;
      if (!hist && keys->len != vals->len) {
        check_keys(c, static_cast<const int64_t *>(dptr(keys)), keys->len, uint64_t(buckets), d_bad);
        read_back(c, d_bad, 8, &bad);
        if (bad != ~uint64_t(0)) bucket_error(c, keys, bad, buckets);
        if (!hist && keys->len != vals->len)
          runtime_error("length mismatch: " + std::to_string(keys->len) + " keys for " +
                        std::to_string(vals->len) + " values");
      }
      vk_array *out = alloc_result(c, d.out_t, uint64_t(buckets));
      if (op == 4) {
        // the keys are checked by the sort inside, before anything is folded
        bad = reduce_into_add_f64(c, static_cast<const int64_t *>(dptr(keys)),
                                  static_cast<const double *>(dptr(vals)), keys->len, uint64_t(buckets),
                                  static_cast<double *>(dptr(out)));
        if (bad != ~uint64_t(0)) {
          release(out);
          bucket_error(c, keys, bad, buckets);
        }
      } else {
        bucket_init(c, op, d.out_t, dptr(out), uint64_t(buckets));
        bucket_rows(c, op, d.out_t, static_cast<const int64_t *>(dptr(keys)), vals ? dptr(vals) : nullptr,
                    keys->len, uint64_t(buckets), dptr(out), d_bad);
        read_back(c, d_bad, 8, &bad);
        if (bad != ~uint64_t(0)) {
          release(out);
          bucket_error(c, keys, bad, buckets);
        }
      }
      v.array = out;
      return v;
    }
    case W_RAND_F64: case W_RAND_I64: {
      if (nargs != 2) arity(2, nargs);
      const uint64_t n = count_arg(args[0]);
      const uint64_t seed = static_cast<uint64_t>(i64_arg(args[1]));
      vk_array *out = alloc_result(c, d.out_t, n);
      if (d.impl == W_RAND_F64) rand_f64(c, n, seed, static_cast<double *>(dptr(out)));
      else rand_i64(c, n, seed, static_cast<int64_t *>(dptr(out)));
      v.array = out;
      return v;
    }
    case W_IOTA: {
      if (nargs != 1) arity(1, nargs);
      const uint64_t n = count_arg(args[0]);
      vk_array *out = alloc_result(c, VK_I64, n);
      iota(c, n, static_cast<int64_t *>(dptr(out)));
      v.array = out;
      return v;
    }
    case W_LENGTH: {
      if (nargs != 1) arity(1, nargs);
      const vk_array *a = array_arg(args[0], d.t0);
      v.kind = VK_VALUE_SCALAR;
      v.scalar = scalar(VK_I64, a->len);
      return v;
    }
    case W_FILL: {
      if (nargs != 2) arity(2, nargs);
      const uint64_t n = count_arg(args[0]);
      if (args[1].kind != VK_SPLAT || args[1].splat.type != d.t0)
        usage_error("expected " + tn(d.t0) + ", got " + arg_ty(args[1]));
      vk_array *out = alloc_result(c, d.t0, n);
      fill_bits(c, type_width(d.t0), n, args[1].splat.bits, dptr(out));
      v.array = out;
      return v;
    }
    default: usage_error("not a whole kernel: " + d.name);
  }
}

void run_elementwise(Ctx &c, const Def &d, const vk_arg *args, size_t nargs, vk_array *out) {
  EwSpec s{};
  s.op = d.p0;
  s.in_type = d.t0;
  s.out_type = d.out_t;
  const void *p[3] = {nullptr, nullptr, nullptr};
  bool *splat[3] = {&s.a_splat, &s.b_splat, &s.c_splat};
  uint64_t *bits[3] = {&s.a_bits, &s.b_bits, &s.c_bits};
  for (size_t i = 0; i < nargs && i < 3; ++i) {
    if (args[i].kind == VK_SPLAT) {
      *splat[i] = true;
      *bits[i] = args[i].splat.bits;
    } else {
      p[i] = dptr(args[i].rows);
    }
  }
  // DivConstBody::run (elementwise.rs:717-736): a zero divisor is the kernel's runtime error
  if (d.impl == E_DIVC && (s.b_bits & (type_width(d.t0) == 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * type_width(d.t0))) - 1)) == 0)
    runtime_error("div_const: division by zero");
  elementwise(c, s, p[0], p[1], p[2], out->len, dptr(out));
}

// Dispatch::call (dispatch.rs:56-141)
vk_value call(Ctx &c, const Def &d, const vk_arg *args, size_t nargs) {
  vk_value v{VK_VALUE_ARRAY, nullptr, {0, 0}};
  if (d.cls == VK_WHOLE) {
    // whole kernels check their arguments themselves; kinds are checked like accepts()
    if (nargs == d.in.size())
      for (size_t i = 0; i < nargs; ++i)
        if (!accepts(d.in[i], args[i])) kind_error(d.in[i], args[i]);
    return run_whole(c, d, args, nargs);
  }
  const int64_t rows = check_args(d, args, nargs);
  switch (d.cls) {
    case VK_ELEMENTWISE: {
      if (rows < 0) usage_error("an array kernel needs an array operand");
      vk_array *out = alloc_result(c, d.out_t, uint64_t(rows));
      run_elementwise(c, d, args, nargs, out);
      v.array = out;
      return v;
    }
    case VK_REDUCTION: {
      if (rows < 0) usage_error("a reduction needs an array operand");
      vk_reducer *r = reducer_new(c, g_by_name[d.name]);
      try {
        if (rows > 0) reducer_update(r, args, nargs);
        v.kind = VK_VALUE_SCALAR;
        v.scalar = reducer_finish(r);
      } catch (...) {
        reducer_free(r);
        throw;
      }
      reducer_free(r);
      return v;
    }
    case VK_SCAN: {
      if (rows < 0) usage_error("a scan needs an array operand");
      vk_array *out = alloc_result(c, d.out_t, uint64_t(rows));
      vk_scanner *s = scanner_new(c, g_by_name[d.name]);
      try {
        if (rows > 0) scanner_update(s, args, nargs, out);
      } catch (...) {
        scanner_free(s);
        release(out);
        throw;
      }
      scanner_free(s);
      v.array = out;
      return v;
    }
    default: {  // selection, one tile with base 0
      vk_builder b{&c, d.out_t, nullptr, 0, 0};
      try {
        select_tile(c, d, args, nargs, 0, &b);
      } catch (...) {
        if (b.ptr) c.pool->free(b.ptr);
        throw;
      }
      v.array = builder_take(&b);
      return v;
    }
  }
}

// ------------------------------------------------------------------ regions
// A fused region (src/exec/region.rs:1-60, region/fused.rs:96-204): elementwise kernels in
// post-order feeding one reduction. When the kernels form one chain over at most two input
// arrays (each kernel reads the previous one's result and a constant, an input array or that
// result again) and the root is a sum, extreme, arg extreme or count, the whole region is one
// device pass with the chain in registers; otherwise it runs kernel by kernel. Both give the
// results of the kernel-by-kernel evaluation bit for bit.

struct RegionPlan {
  ChainSpec ch{};
  const vk_array *x = nullptr, *y = nullptr;
  const Def *root = nullptr;
  ReduceSpec rs{};
  bool canon = false;
};

bool same_view(const vk_array *a, const vk_array *b) {
  return a == b || (a->buf == b->buf && a->offset == b->offset && a->len == b->len && a->type == b->type);
}

int cmp_swapped(int cmp) {  // c <cmp> v as v <cmp'> c
  switch (cmp) {
    case CMP_LT: return CMP_GT;
    case CMP_LE: return CMP_GE;
    case CMP_GT: return CMP_LT;
    case CMP_GE: return CMP_LE;
    default: return cmp;
  }
}

bool region_compile(const vk_rkernel *ks, size_t nk, const char *root, const vk_src *rargs, size_t nr,
                    const vk_arg *in, size_t nin, RegionPlan &p) {
  build_table();
  auto rit = g_by_name.find(root ? root : "");
  if (rit == g_by_name.end() || nk == 0 || nk > size_t(kChainMax) + 1) return false;
  const Def &rd = g_defs[rit->second];
  if (rd.cls != VK_REDUCTION) return false;
  const bool root_ok = (rd.impl == R_CANON && rd.p0 == 0) || rd.impl == R_SUM_INT || rd.impl == R_EXTREME ||
                       rd.impl == R_ARG || rd.impl == R_COUNT_CMP || rd.impl == R_COUNT_BOOL;
  if (!root_ok || nr != rd.in.size()) return false;
  if (rargs[0].kind != VK_SRC_KERNEL || rargs[0].index != int32_t(nk) - 1) return false;
  for (size_t i = 1; i < nr; ++i) {  // count_<cmp>_<t>'s constant
    if (rargs[i].kind != VK_SRC_INPUT || rargs[i].index < 0 || size_t(rargs[i].index) >= nin) return false;
    if (!accepts(rd.in[i], in[rargs[i].index]) || in[rargs[i].index].kind != VK_SPLAT) return false;
  }
  int32_t T = -1;
  int nops = 0;
  bool count_cmp = false;
  for (size_t j = 0; j < nk; ++j) {
    auto it = g_by_name.find(ks[j].name ? ks[j].name : "");
    if (it == g_by_name.end()) return false;
    const Def &d = g_defs[it->second];
    if (d.cls != VK_ELEMENTWISE || ks[j].nargs < 1 || ks[j].nargs > 3 || size_t(ks[j].nargs) != d.in.size()) return false;
    const bool cmp = d.impl == E_BINARY && d.p0 >= EW_EQ && d.p0 <= EW_GE;
    const bool arith = (d.impl == E_BINARY && (d.p0 == EW_ADD || d.p0 == EW_SUB || d.p0 == EW_MUL || d.p0 == EW_DIV ||
                                               d.p0 == EW_MIN || d.p0 == EW_MAX || d.p0 == EW_BAND ||
                                               d.p0 == EW_BOR || d.p0 == EW_BXOR)) ||
                       (d.impl == E_UNARY && (d.p0 == EW_NEG || d.p0 == EW_ABS || d.p0 == EW_SQRT || d.p0 == EW_BNOT));
    if (!(arith || (cmp && j + 1 == nk && rd.impl == R_COUNT_BOOL))) return false;
    if (d.t0 == VK_BOOL) return false;
    if (T < 0) T = d.t0;
    if (d.t0 != T || (!cmp && d.out_t != T)) return false;
    // classify the arguments: the chain value, a constant, or an input array
    int chain_at = -1, other_at = -1;
    bool self = false;
    for (int a = 0; a < ks[j].nargs; ++a) {
      const vk_src &sr = ks[j].args[a];
      if (sr.kind == VK_SRC_KERNEL) {
        if (j == 0 || sr.index != int32_t(j) - 1) return false;
        if (chain_at >= 0) self = true;
        else chain_at = a;
        continue;
      }
      if (sr.kind != VK_SRC_INPUT || sr.index < 0 || size_t(sr.index) >= nin) return false;
      if (!accepts(d.in[a], in[sr.index])) return false;
    }
    ChainOp op{};
    if (j == 0) {
      // the chain starts at the first array argument, which becomes x
      for (int a = 0; a < ks[j].nargs && chain_at < 0; ++a)
        if (in[ks[j].args[a].index].kind == VK_ROWS) chain_at = a;
      if (chain_at < 0) return false;
      p.x = in[ks[j].args[chain_at].index].rows;
    } else if (chain_at < 0) {
      return false;
    }
    if (d.in.size() == 2) other_at = 1 - chain_at;
    if (self) {
      op.src = 3;
    } else if (other_at >= 0) {
      const vk_arg &o = in[ks[j].args[other_at].index];
      if (o.kind == VK_SPLAT) {
        op.src = 0;
        op.kbits = o.splat.bits;
      } else if (same_view(o.rows, p.x)) {
        op.src = 1;
      } else if (!p.y || same_view(o.rows, p.y)) {
        p.y = o.rows;
        op.src = 2;
      } else {
        return false;
      }
    }
    op.swap = chain_at == 1 ? 1 : 0;
    if (cmp) {
      // count(<cmp>(chain, c)): the root counts the chain's rows that pass the comparison
      if (op.src != 0) return false;
      p.rs = ReduceSpec{};
      p.rs.kind = ReduceSpec::COUNT_CMP;
      p.rs.type = T;
      p.rs.cmp = op.swap ? cmp_swapped(d.p0 - EW_EQ) : d.p0 - EW_EQ;
      p.rs.cbits = op.kbits;
      count_cmp = true;
      continue;
    }
    op.op = d.p0;
    if (nops >= kChainMax) return false;
    p.ch.ops[nops++] = op;
  }
  p.ch.n = nops;
  p.ch.type = T;
  if (p.y && p.y->len != p.x->len) return false;
  p.root = &rd;
  if (rd.impl == R_COUNT_BOOL) {
    if (!count_cmp) return false;
  } else {
    if (rd.t0 != T) return false;
    if (rd.impl == R_CANON) {
      p.canon = true;
      return T == VK_F64;
    }
    p.rs = reduce_spec(rd, 0);
    if (rd.impl == R_COUNT_CMP) p.rs.cbits = in[rargs[1].index].splat.bits;
  }
  return p.y == nullptr;  // the streaming reductions read one input
}

// The kernel-by-kernel evaluation of a region: every kernel as a whole call, then the root.
vk_scalar region_eval(Ctx &c, const vk_rkernel *ks, size_t nk, const char *root, const vk_src *rargs, size_t nr,
                      const vk_arg *in, size_t nin) {
  std::vector<vk_value> vals;
  auto release_all = [&] {
    for (auto &v : vals)
      if (v.kind == VK_VALUE_ARRAY && v.array) release(v.array);
  };
  auto arg_of = [&](const vk_src &sr, size_t upto) -> vk_arg {
    if (sr.kind == VK_SRC_INPUT) {
      if (sr.index < 0 || size_t(sr.index) >= nin) usage_error("region input " + std::to_string(sr.index) + " out of range");
      return in[sr.index];
    }
    if (sr.kind != VK_SRC_KERNEL || sr.index < 0 || size_t(sr.index) >= upto)
      usage_error("region kernel " + std::to_string(sr.index) + " is not an earlier kernel");
    const vk_value &v = vals[sr.index];
    vk_arg a{};
    if (v.kind == VK_VALUE_ARRAY) {
      a.kind = VK_ROWS;
      a.rows = v.array;
    } else {
      a.kind = VK_SPLAT;
      a.splat = v.scalar;
    }
    return a;
  };
  try {
    for (size_t j = 0; j < nk; ++j) {
      const Def &d = def_of(ks[j].name);
      if (d.cls != VK_ELEMENTWISE) usage_error(d.name + " is not an elementwise kernel");
      std::vector<vk_arg> a;
      for (int k = 0; k < ks[j].nargs && k < 3; ++k) a.push_back(arg_of(ks[j].args[k], j));
      vals.push_back(call(c, d, a.data(), a.size()));
    }
    const Def &rd = def_of(root);
    if (rd.cls != VK_REDUCTION) usage_error(rd.name + " is not a reduction");
    std::vector<vk_arg> a;
    for (size_t k = 0; k < nr; ++k) a.push_back(arg_of(rargs[k], nk));
    vk_value r = call(c, rd, a.data(), a.size());
    release_all();
    return r.scalar;
  } catch (...) {
    release_all();
    throw;
  }
}

vk_scalar region_run(Ctx &c, const RegionPlan &p) {
  const uint64_t n = p.x->len;
  const void *x = dptr(p.x);
  const void *y = p.y ? dptr(p.y) : nullptr;
  void *state = tmp_alloc(c, 256);
  vk_scalar out{};
  try {
    if (p.canon) {
      double *acc = static_cast<double *>(state);
      canon_init(c, false, acc);
      const uint64_t blocks = n / 16, rest = n % 16;
      const double *xd = static_cast<const double *>(x), *yd = static_cast<const double *>(y);
      canon_blocks_chain(c, p.ch, xd, yd, blocks, acc);
      canon_rest_chain(c, p.ch, xd + blocks * 16, yd ? yd + blocks * 16 : nullptr, rest, acc + 16);
      canon_finish(c, false, acc, acc + 16, static_cast<int>(rest), reinterpret_cast<double *>(c.scratch.slots));
      uint64_t b;
      read_back(c, c.scratch.slots, 8, &b);
      out = scalar(VK_F64, b);
    } else {
      reduce_state_init(c, p.rs, state);
      reduce_rows_chain(c, p.rs, p.ch, x, n, state);
      uint64_t w[3];
      read_back(c, state, 24, w);
      out = finish_words(*p.root, w);
    }
  } catch (...) {
    tmp_free(c, state);
    throw;
  }
  tmp_free(c, state);
  return out;
}

// ------------------------------------------------------------------ batching

std::vector<vk_value> batch_call_all(Ctx &c, const Def &d, const vk_arg *args, size_t nargs, size_t nitems);

// Items of at least this many rows run as single calls in a batched argsort, group_id or
// reduce_into_add_f64: the batched forms pay for their per-item tables and the single-call
// kernels are faster on long inputs (4096 arrays of 1K to 16M rows ran 1.2x slower batched).
constexpr uint64_t kBatchLargeRows = uint64_t(1) << 22;

void release_values(std::vector<vk_value> &vs) {
  for (auto &v : vs)
    if (v.kind == VK_VALUE_ARRAY && v.array) {
      release(v.array);
      v.array = nullptr;
    }
}

std::vector<vk_value> batch_call(Ctx &c, const Def &d, const vk_arg *args, size_t nargs, size_t nitems) {
  const bool split = d.cls == VK_WHOLE && nargs >= 1 &&
                     (d.impl == W_ARGSORT || d.impl == W_GROUP_ID || (d.impl == W_REDUCE_INTO && d.p0 == 4));
  std::vector<size_t> small, large;
  if (split)
    for (size_t i = 0; i < nitems; ++i) {
      const vk_arg &a = args[i * nargs];
      const uint64_t n = a.kind == VK_ROWS && a.rows ? a.rows->len : 0;
      (n >= kBatchLargeRows ? large : small).push_back(i);
    }
  if (large.empty()) return batch_call_all(c, d, args, nargs, nitems);
  std::vector<vk_value> out(nitems, vk_value{VK_VALUE_ARRAY, nullptr, {0, 0}});
  try {
    for (size_t i : large) out[i] = call(c, d, args + i * nargs, nargs);
    if (!small.empty()) {
      std::vector<vk_arg> sub;
      sub.reserve(small.size() * nargs);
      for (size_t i : small)
        for (size_t k = 0; k < nargs; ++k) sub.push_back(args[i * nargs + k]);
      std::vector<vk_value> r = batch_call_all(c, d, sub.data(), nargs, small.size());
      for (size_t j = 0; j < small.size(); ++j) out[small[j]] = r[j];
    }
  } catch (...) {
    // the first failing item's error, as one call per item in order reports it
    release_values(out);
    for (size_t i = 0; i < nitems; ++i) {
      try {
        out[i] = call(c, d, args + i * nargs, nargs);
      } catch (...) {
        release_values(out);
        throw;
      }
    }
  }
  return out;
}

std::vector<vk_value> batch_call_all(Ctx &c, const Def &d, const vk_arg *args, size_t nargs, size_t nitems) {
  std::vector<vk_value> out(nitems, vk_value{VK_VALUE_ARRAY, nullptr, {0, 0}});
  // every item's arguments pass the checks of a single call first
  std::vector<int64_t> rows(nitems, -1);
  for (size_t i = 0; i < nitems; ++i) {
    const vk_arg *a = args + i * nargs;
    if (d.cls == VK_WHOLE) {
      if (nargs != d.in.size()) arity(d.in.size(), nargs);
      for (size_t k = 0; k < nargs; ++k)
        if (!accepts(d.in[k], a[k])) kind_error(d.in[k], a[k]);
    } else {
      rows[i] = check_args(d, a, nargs);
    }
  }
  auto each = [&]() {
    size_t done = 0;
    try {
      for (; done < nitems; ++done) out[done] = call(c, d, args + done * nargs, nargs);
    } catch (...) {
      for (size_t i = 0; i < done; ++i) vk_value_release(&out[i]);
      throw;
    }
  };
  for (size_t i = 0; i < nitems; ++i)
    if (d.cls != VK_WHOLE && rows[i] < 0) {  // no array operand: the single call's own error
      each();
      return out;
    }
  if (d.cls == VK_SELECTION) {
    const uint32_t w = type_width(d.out_t);
    uint64_t total = 0;
    std::vector<uint64_t> off(nitems);
    for (size_t i = 0; i < nitems; ++i) {
      off[i] = total;
      total += rows[i] < 0 ? 0 : uint64_t(rows[i]);
    }
    vk_array *buf = new_array(c, d.out_t, total ? total : 1);
    std::vector<BatchItem> items(nitems);
    SelectSpec s{};
    for (size_t i = 0; i < nitems; ++i) {
      const vk_arg *a = args + i * nargs;
      BatchItem &it = items[i];
      it.n = rows[i] < 0 ? 0 : uint64_t(rows[i]);
      it.out = static_cast<char *>(dptr(buf)) + off[i] * w;
      if (d.impl == S_WHERE) {
        it.a = dptr(a[0].rows);
        it.aux = 0;
      } else {
        it.a = dptr(a[0].rows);
        if (d.impl == S_COMPRESS) it.b = dptr(a[1].rows);
      }
    }
    if (d.impl == S_COMPRESS) {
      s.kind = 0;
      s.type = d.t0;
    } else if (d.impl == S_FILTER) {
      s.kind = 1;
      s.type = d.t0;
      s.cmp = d.p0;
      s.cbits = args[1].splat.bits;
      for (size_t i = 1; i < nitems; ++i)
        if (args[i * nargs + 1].splat.bits != s.cbits) {  // one constant per launch
          release(buf);
          each();
          return out;
        }
    } else {
      s.kind = 2;
      s.type = VK_I64;
    }
    auto *counts = static_cast<uint64_t *>(tmp_alloc(c, nitems * 8));
    select_batch(c, s, items, counts);
    std::vector<uint64_t> got(nitems);
    check_hip(hipMemcpyAsync(got.data(), counts, nitems * 8, hipMemcpyDeviceToHost, c.stream), "counts");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    tmp_free(c, counts);
    for (size_t i = 0; i < nitems; ++i) {
      buf->buf->refs++;
      out[i].array = new vk_array{1, buf->buf, buf->offset + off[i] * w, got[i], d.out_t};
    }
    release(buf);
    return out;
  }
  if (d.impl == R_FSUM_INT || (d.impl == R_CANON && (d.p0 == 4 || d.t0 != VK_F64))) {
    // no batched form: integer dot/sumsq/condsum and the f32 canonical folds run item by item
    each();
    return out;
  }
  if (d.impl == R_CANON && d.p0 != 4) {
    std::vector<BatchItem> items(nitems);
    CanonSpec cs{d.p0, d.p1, 0, 0};
    for (size_t i = 0; i < nitems; ++i) {
      const vk_arg *a = args + i * nargs;
      items[i].a = dptr(a[0].rows);
      items[i].b = d.p0 == 1 ? dptr(a[1].rows) : nullptr;
      items[i].n = uint64_t(rows[i]);
      if (d.p0 == 3) {
        if (a[1].splat.bits != args[1].splat.bits || a[2].splat.bits != args[2].splat.bits) {
          each();
          return out;
        }
      }
    }
    if (d.p0 == 3) {
      std::memcpy(&cs.c, &args[1].splat.bits, 8);
      std::memcpy(&cs.e, &args[2].splat.bits, 8);
    }
    auto *res = static_cast<double *>(tmp_alloc(c, nitems * 8));
    canon_batch(c, cs, items, res);
    std::vector<uint64_t> bits(nitems);
    check_hip(hipMemcpyAsync(bits.data(), res, nitems * 8, hipMemcpyDeviceToHost, c.stream), "results");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    tmp_free(c, res);
    for (size_t i = 0; i < nitems; ++i) {
      out[i].kind = VK_VALUE_SCALAR;
      out[i].scalar = scalar(VK_F64, bits[i]);
    }
    return out;
  }
  if (d.cls == VK_REDUCTION && d.impl != R_CANON) {
    // the other reductions: one fold of every item's tiles, one merge per item
    ReduceSpec rs = reduce_spec(d, 0);
    if (d.impl == R_COUNT_CMP) {
      rs.cbits = splat_arg(args[1], d.t0);
      for (size_t i = 1; i < nitems; ++i)
        if (args[i * nargs + 1].splat.bits != rs.cbits) {  // one constant per launch
          each();
          return out;
        }
    }
    std::vector<BatchItem> items(nitems);
    for (size_t i = 0; i < nitems; ++i) {
      const vk_array *x = args[i * nargs].rows;
      items[i].a = dptr(x);
      items[i].n = x->len;
    }
    const uint32_t words = reduce_state_words(rs);
    auto *states = static_cast<uint64_t *>(tmp_alloc(c, nitems * words * 8));
    std::vector<uint64_t> host(nitems * words);
    try {
      reduce_batch(c, rs, items, states);
      check_hip(hipMemcpyAsync(host.data(), states, nitems * words * 8, hipMemcpyDeviceToHost, c.stream),
                "states");
      check_hip(hipStreamSynchronize(c.stream), "sync");
    } catch (...) {
      tmp_free(c, states);
      throw;
    }
    tmp_free(c, states);
    for (size_t i = 0; i < nitems; ++i) {
      uint64_t w[3] = {0, 0, 0};
      for (uint32_t k = 0; k < words; ++k) w[k] = host[i * words + k];
      out[i].kind = VK_VALUE_SCALAR;
      out[i].scalar = finish_words(d, w);
    }
    return out;
  }
  if (d.impl == W_GROUP_ID) {
    std::vector<const vk_array *> xs(nitems);
    std::vector<uint64_t> off(nitems);
    uint64_t total = 0;
    for (size_t i = 0; i < nitems; ++i) {
      xs[i] = array_arg(args[i * nargs], d.t0);
      off[i] = total;
      total += xs[i]->len;
    }
    vk_array *buf = new_array(c, VK_I64, total ? total : 1);
    std::vector<BatchItem> items(nitems);
    for (size_t i = 0; i < nitems; ++i) {
      items[i].a = dptr(xs[i]);
      items[i].n = xs[i]->len;
      items[i].out = static_cast<int64_t *>(dptr(buf)) + off[i];
    }
    auto *groups = static_cast<uint64_t *>(tmp_alloc(c, nitems * 8));
    try {
      group_id_batch(c, d.t0, items, groups);
    } catch (...) {
      tmp_free(c, groups);
      release(buf);
      throw;
    }
    tmp_free(c, groups);
    for (size_t i = 0; i < nitems; ++i) {
      buf->buf->refs++;
      out[i].array = new vk_array{1, buf->buf, buf->offset + off[i] * 8, xs[i]->len, VK_I64};
    }
    release(buf);
    return out;
  }
  if (d.impl == W_REDUCE_INTO && d.p0 == 4) {
    // host-side errors (bucket count, lengths) keep the single call's order item by item
    std::vector<BatchItem> items(nitems);
    std::vector<const vk_array *> ks(nitems);
    std::vector<uint64_t> off(nitems);
    uint64_t total = 0;
    for (size_t i = 0; i < nitems; ++i) {
      const vk_arg *a = args + i * nargs;
      const vk_array *vals = array_arg(a[1], VK_F64);
      ks[i] = array_arg(a[0], VK_I64);
      const int64_t nb = i64_arg(a[2]);
      if (nb < 0 || ks[i]->len != vals->len) {
        each();
        return out;
      }
      items[i].a = dptr(ks[i]);
      items[i].b = dptr(vals);
      items[i].n = ks[i]->len;
      items[i].aux2 = static_cast<uint64_t>(nb);
      off[i] = total;
      total += static_cast<uint64_t>(nb);
    }
    auto *bad = static_cast<uint64_t *>(tmp_alloc(c, nitems * 8));
    std::vector<uint64_t> hbad(nitems);
    check_keys_batch(c, items, bad);
    check_hip(hipMemcpyAsync(hbad.data(), bad, nitems * 8, hipMemcpyDeviceToHost, c.stream), "bad keys");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    tmp_free(c, bad);
    for (size_t i = 0; i < nitems; ++i)
      if (hbad[i] != ~uint64_t(0)) bucket_error(c, ks[i], hbad[i], static_cast<int64_t>(items[i].aux2));
    vk_array *buf = new_array(c, VK_F64, total ? total : 1);
    for (size_t i = 0; i < nitems; ++i) items[i].out = static_cast<double *>(dptr(buf)) + off[i];
    try {
      reduce_into_add_f64_batch(c, items);
    } catch (...) {
      release(buf);
      throw;
    }
    for (size_t i = 0; i < nitems; ++i) {
      buf->buf->refs++;
      out[i].array = new vk_array{1, buf->buf, buf->offset + off[i] * 8, items[i].aux2, VK_F64};
    }
    release(buf);
    return out;
  }
  if (d.impl == W_ARGSORT) {
    std::vector<const vk_array *> xs(nitems);
    std::vector<uint64_t> off(nitems);
    uint64_t total = 0;
    for (size_t i = 0; i < nitems; ++i) {
      xs[i] = array_arg(args[i * nargs], d.t0);
      if (xs[i]->len > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
      off[i] = total;
      total += xs[i]->len;
    }
    vk_array *buf = new_array(c, VK_I64, total ? total : 1);
    std::vector<BatchItem> items(nitems);
    for (size_t i = 0; i < nitems; ++i) {
      items[i].a = dptr(xs[i]);
      items[i].n = xs[i]->len;
      items[i].out = static_cast<int64_t *>(dptr(buf)) + off[i];
    }
    try {
      argsort_batch(c, d.t0, items);
    } catch (...) {
      release(buf);
      throw;
    }
    for (size_t i = 0; i < nitems; ++i) {
      buf->buf->refs++;
      out[i].array = new vk_array{1, buf->buf, buf->offset + off[i] * 8, xs[i]->len, VK_I64};
    }
    release(buf);
    return out;
  }
  each();
  return out;
}

int set_status(vk_status *st, int code, const std::string &msg) {
  if (st) {
    st->code = code;
    std::snprintf(st->msg, sizeof st->msg, "%s", msg.c_str());
  }
  return code;
}

template <typename F>
int guarded(vk_status *st, F &&f) {
  try {
    f();
    return set_status(st, VK_OK, "");
  } catch (const Error &e) {
    return set_status(st, e.code, e.msg);
  } catch (const std::exception &e) {
    return set_status(st, VK_ERR_DEVICE, e.what());
  }
}

// ------------------------------------------------------------- batch plans
// A plan of batched statements over a fixed batch of resident arrays (vk_bplan_*): every
// statement is one launch sequence over all items, its tiles planned on the device from the
// row counts earlier statements left there (a selection's kept counts), so the whole plan
// runs with no host read-back, as one sequence of kernels with arguments fixed at compile time,
// captured once into a hipGraph and replayed. Statements:
// selections (filter_<cmp>_<t> over a plan input or an earlier selection, compress_<t> over
// plan inputs), canonical float sums (reduce_add_f64, sumsq_f64, condsum_<cmp>_f64 over a plan
// input or a selection; dot_f64 over plan inputs) and the other reductions (over a plan input
// or a selection). Argument references: a plan input (per item), an earlier statement's
// per-item array, or a scalar.

__global__ void k_plan_bump(uint32_t *epoch, uint32_t *wrapped) {
  const uint32_t e = *epoch;
  const uint32_t n = e >= 65535u ? 1u : e + 1u;
  *epoch = n;
  *wrapped = n == 1u ? 1u : 0u;
}

// Clears look-back status words when the plan epoch has wrapped (never in between), so a
// word from 65535 runs ago cannot pass for this run's.
__global__ void k_plan_clear(uint64_t *status, uint64_t words, const uint32_t *wrapped) {
  if (!*wrapped) return;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < words; i += uint64_t(gridDim.x) * blockDim.x)
    status[i] = 0;
}

// Compiling a plan captures its launch sequence (kernels only, arguments fixed at compile time)
// into a hipGraph; a run replays the graph.
constexpr bool kCapturePlans = true;

struct PlanStmt {
  const Def *d = nullptr;
  int kind = 0;  // 0 selection, 1 canonical sum, 2 other reduction, 3 argsort, 4 group_id,
                 // 5 reduce_into_add_f64 over a group_id statement's ids and group counts
  std::vector<vk_bref> args;
  std::vector<uint64_t> bound;  // per item: an upper bound of its rows
  int src = -1;                 // statement whose arrays are the rows (or -1: plan inputs)
  DevPlan dp;
  uint64_t status_words = 0;
  uint64_t *counts = nullptr;   // selection: kept rows per item
  vk_array *outbuf = nullptr;   // selection: the kept rows of item i from row off[i]
  std::vector<uint64_t> off;
  uint64_t *res = nullptr;      // reductions: nitems x words result words
  double *acc = nullptr, *zero = nullptr;
  void *partials = nullptr;
  uint32_t words = 1;
  SelectSpec ss{};
  CanonSpec cs{};
  ReduceSpec rs{};
  int items_per_thread = 16;    // selection rows per thread (autotuned: 8, 16 or 0, the persistent filter)
  uint64_t tile_rows = 8192;    // reduction rows per tile (autotuned: 4096, 8192 or 16384)
  std::vector<uint64_t> host;   // results copied back after a run
  int gsrc = -1;                // reduce_into_add_f64: the group_id statement
  PlanSort *sort = nullptr;     // argsort
  PlanGroup *group = nullptr;   // group_id (its group counts at group->groups)
  PlanFold *fold = nullptr;     // reduce_into_add_f64
  std::vector<std::pair<uint64_t *, uint64_t>> clear;  // look-back status words (cleared on epoch wrap)
};

}  // namespace
}  // namespace vk

struct vk_bplan {
  vk::Ctx *c = nullptr;
  size_t nitems = 0, ninputs = 0;
  std::vector<vk_array *> inputs;  // item-major, retained
  std::vector<vk::PlanStmt> st;
  uint32_t *epoch = nullptr;       // [0] epoch, [1] wrapped flag
  std::vector<void *> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool compiled = false, fresh = false;
  std::vector<std::string> tuned;  // autotuner decisions, one line per tuned statement
};

namespace vk {
namespace {

void *plan_alloc(vk_bplan &p, uint64_t bytes) {
  void *q = tmp_alloc(*p.c, bytes ? bytes : 8);
  p.allocs.push_back(q);
  return q;
}

// The rows of argument `a` of item i: pointer (null for a selection not compiled yet), type
// and bound (rows known on the host or an upper bound when they come from a selection's device
// count).
struct RowsRef {
  const void *ptr;
  int32_t type;
  uint64_t bound;
  int src;  // -1 plan input, else the selection statement
};

RowsRef rows_ref(const vk_bplan &p, const vk_bref &a, size_t i) {
  if (a.kind == VK_BREF_INPUT) {
    if (a.index < 0 || size_t(a.index) >= p.ninputs) usage_error("plan input out of range");
    const vk_array *x = p.inputs[i * p.ninputs + size_t(a.index)];
    return RowsRef{dptr(x), x->type, x->len, -1};
  }
  if (a.kind == VK_BREF_STMT) {
    if (a.index < 0 || size_t(a.index) >= p.st.size()) usage_error("plan statement out of range");
    const PlanStmt &s = p.st[size_t(a.index)];
    if (s.kind == 1 || s.kind == 2) usage_error("statement " + std::to_string(a.index) + " has no array result");
    // the statement's buffer exists once it is compiled (compiling goes in statement order)
    const uint32_t w = type_width(s.d->out_t);
    const void *ptr = s.outbuf ? static_cast<const char *>(dptr(s.outbuf)) + s.off[i] * w : nullptr;
    // argsort and group_id results have the rows of their input; selections and bucket sums
    // have device row counts
    const int src = (s.kind == 3 || s.kind == 4) ? -1 : a.index;
    return RowsRef{ptr, s.d->out_t, s.bound[i], src};
  }
  usage_error("a plan argument of " + std::string("kind ") + std::to_string(a.kind) + " is not rows");
}

uint64_t tiles_for(uint64_t rows, uint64_t unit, uint64_t per) { return (rows / unit + per - 1) / per; }

// Statement s's launches over plan state (capturable: kernels only).
void plan_launch(vk_bplan &p, PlanStmt &s) {
  Ctx &c = *p.c;
  const uint32_t ni = static_cast<uint32_t>(p.nitems);
  const uint64_t *dev_n = s.src >= 0 ? p.st[size_t(s.src)].counts : nullptr;
  if (s.kind == 0) {
    plan_tiles_dev(c, s.dp, ni, dev_n, 1, select_tile_rows(s.items_per_thread));
    select_batch_dev(c, s.ss, s.items_per_thread, s.dp, ni, s.counts);
  } else if (s.kind == 1) {
    plan_tiles_dev(c, s.dp, ni, dev_n, 16, canon_tile_blocks(), canon_tile_min_rows());
    canon_batch_dev(c, s.cs, s.dp, ni, s.acc, s.zero, reinterpret_cast<double *>(s.res));
  } else if (s.kind == 2) {
    plan_tiles_dev(c, s.dp, ni, dev_n, 1, s.tile_rows);
    reduce_batch_dev(c, s.rs, s.tile_rows, s.dp, ni, s.partials, s.res);
  } else if (s.kind == 3) {
    plan_sort_run(c, *s.sort, p.epoch);
  } else if (s.kind == 4) {
    plan_group_run(c, *s.group);
  } else {
    plan_fold_run(c, *s.fold, p.epoch);
  }
}

void plan_prologue(vk_bplan &p) {
  Ctx &c = *p.c;
  k_plan_bump<<<1, 1, 0, c.stream>>>(p.epoch, p.epoch + 1);
  for (auto &s : p.st)
    for (auto &r : s.clear)
      if (r.second)
        k_plan_clear<<<stream_grid(r.second, kBlock, 256), kBlock, 0, c.stream>>>(r.first, r.second, p.epoch + 1);
  check_hip(hipGetLastError(), "plan prologue");
}

void plan_run_eager(vk_bplan &p) {
  plan_prologue(p);
  for (auto &s : p.st) plan_launch(p, s);
}

// Autotuning: each tunable statement is timed with every candidate on the plan's own inputs
// (its sources already computed by an eager run), best of three launches each.
void plan_tune(vk_bplan &p) {
  Ctx &c = *p.c;
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "event");
  check_hip(hipEventCreate(&e1), "event");
  auto time_stmt = [&](PlanStmt &s) {
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      plan_prologue(p);  // a new epoch for every timed launch of a look-back kernel
      check_hip(hipEventRecord(e0, c.stream), "event");
      plan_launch(p, s);
      check_hip(hipEventRecord(e1, c.stream), "event");
      check_hip(hipEventSynchronize(e1), "event");
      float ms = 0;
      check_hip(hipEventElapsedTime(&ms, e0, e1), "event");
      if (ms < best) best = ms;
    }
    return best;
  };
  plan_run_eager(p);
  for (size_t k = 0; k < p.st.size(); ++k) {
    PlanStmt &s = p.st[k];
    std::string line = "statement " + std::to_string(k) + " " + s.d->name + ":";
    if (s.kind == 0) {
      // one tile per block with 8 or 16 rows per thread, or (filters) the persistent form: one
      // block per compute unit loading its next tile during each look-back
      int best = 16;
      float bt = 1e30f;
      std::vector<int> cands = {8, 16};
      if (s.ss.kind == 1) cands.push_back(0);
      for (int ipt : cands) {
        s.items_per_thread = ipt;
        const float t = time_stmt(s);
        line += ipt ? " " + std::to_string(ipt) + " rows/thread " : std::string(" persistent ");
        line += std::to_string(t * 1000.0f) + " us;";
        if (t < bt) { bt = t; best = ipt; }
      }
      s.items_per_thread = best;
      line += best ? " chose " + std::to_string(best) : std::string(" chose persistent");
    } else if (s.kind == 2) {
      uint64_t best = 8192;
      float bt = 1e30f;
      for (uint64_t tr : {uint64_t(4096), uint64_t(8192), uint64_t(16384)}) {
        s.tile_rows = tr;
        const float t = time_stmt(s);
        line += " " + std::to_string(tr) + "-row tiles " + std::to_string(t * 1000.0f) + " us;";
        if (t < bt) { bt = t; best = tr; }
      }
      s.tile_rows = best;
      line += " chose " + std::to_string(best);
    } else {
      continue;  // canonical sums have one tile shape (16K rows, the exact form's)
    }
    p.tuned.push_back(line);
    // later statements read this one's outputs: recompute them with the choice
    plan_prologue(p);
    plan_launch(p, s);
  }
  check_hip(hipStreamSynchronize(c.stream), "tune");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void plan_add(vk_bplan &p, const char *name, const vk_bref *args, size_t nargs) {
  if (p.compiled) usage_error("the plan is already compiled");
  const Def &d = def_of(name);
  PlanStmt s;
  s.d = &d;
  s.args.assign(args, args + nargs);
  if (nargs != d.in.size()) arity(d.in.size(), nargs);
  auto rows_type = [&](size_t k) {
    if (args[k].kind == VK_BREF_SCALAR) usage_error("expected " + tn(d.in[k].t) + "[], got " + tn(args[k].scalar.type));
    const RowsRef r = rows_ref(p, args[k], 0);
    if (r.type != d.in[k].t) usage_error("expected " + tn(d.in[k].t) + "[], got " + tn(r.type) + "[]");
    return r;
  };
  auto scalar_of = [&](size_t k) {
    if (args[k].kind != VK_BREF_SCALAR || args[k].scalar.type != d.in[k].t)
      usage_error("expected " + tn(d.in[k].t) + " as argument " + std::to_string(k));
    return args[k].scalar.bits;
  };
  if (d.cls == VK_SELECTION && (d.impl == S_FILTER || d.impl == S_COMPRESS)) {
    s.kind = 0;
    const RowsRef r0 = rows_type(0);
    s.src = r0.src;
    if (d.impl == S_FILTER) {
      s.ss.kind = 1;
      s.ss.type = d.t0;
      s.ss.cmp = d.p0;
      s.ss.cbits = scalar_of(1);
    } else {
      const RowsRef r1 = rows_type(1);
      if (r0.src >= 0 || r1.src >= 0) usage_error("a planned compress takes plan inputs");
      s.ss.kind = 0;
      s.ss.type = d.t0;
    }
  } else if (d.cls == VK_REDUCTION && (d.impl == R_FSUM_INT || (d.impl == R_CANON && d.t0 != VK_F64))) {
    usage_error(d.name + " cannot be planned (f64 canonical sums, selections and the other reductions can)");
  } else if (d.cls == VK_REDUCTION && d.impl == R_CANON && d.p0 != 4) {
    s.kind = 1;
    const RowsRef r0 = rows_type(0);
    s.src = r0.src;
    s.cs = CanonSpec{d.p0, d.p1, 0, 0};
    if (d.p0 == 1) {
      const RowsRef r1 = rows_type(1);
      if (r0.src >= 0 || r1.src >= 0) usage_error("a planned dot_f64 takes plan inputs");
    }
    if (d.p0 == 3) {
      const uint64_t cb = scalar_of(1), eb = scalar_of(2);
      std::memcpy(&s.cs.c, &cb, 8);
      std::memcpy(&s.cs.e, &eb, 8);
    }
  } else if (d.cls == VK_REDUCTION && d.impl != R_CANON) {
    s.kind = 2;
    const RowsRef r0 = rows_type(0);
    s.src = r0.src;
    s.rs = reduce_spec(d, 0);
    if (d.impl == R_COUNT_CMP) s.rs.cbits = scalar_of(1);
    s.words = reduce_state_words(s.rs);
  } else if (d.cls == VK_WHOLE && (d.impl == W_ARGSORT || d.impl == W_GROUP_ID)) {
    s.kind = d.impl == W_ARGSORT ? 3 : 4;
    const RowsRef r0 = rows_type(0);
    if (r0.src >= 0) usage_error("a planned " + d.name + " takes a plan input or an argsort or group_id result");
  } else if (d.cls == VK_WHOLE && d.impl == W_REDUCE_INTO && d.p0 == 4) {
    // reduce_into_add_f64(ids, values, buckets) with ids and buckets from one group_id statement
    s.kind = 5;
    auto group_stmt = [&](size_t k) {
      if (args[k].kind != VK_BREF_STMT || args[k].index < 0 || size_t(args[k].index) >= p.st.size() ||
          p.st[size_t(args[k].index)].kind != 4)
        usage_error("a planned reduce_into_add_f64 takes the ids and the group count of a group_id statement");
      return args[k].index;
    };
    s.gsrc = group_stmt(0);
    if (group_stmt(2) != s.gsrc) usage_error("the ids and the bucket count come from different statements");
    const RowsRef r1 = rows_type(1);
    if (r1.src >= 0) usage_error("a planned reduce_into_add_f64 takes plan input values");
  } else {
    usage_error(d.name + " cannot be planned (selections, canonical sums, reductions, argsort, group_id and "
                "reduce_into_add_f64 over a group_id can)");
  }
  // per-item bounds and equal lengths of two-array statements
  s.bound.resize(p.nitems);
  for (size_t i = 0; i < p.nitems; ++i) {
    const RowsRef r0 = rows_ref(p, s.args[0], i);
    s.bound[i] = r0.bound;
    if ((s.kind == 0 && d.impl == S_COMPRESS) || (s.kind == 1 && d.p0 == 1) || s.kind == 5) {
      const RowsRef r1 = rows_ref(p, s.args[1], i);
      if (r1.bound != r0.bound)
        usage_error("length mismatch: " + std::to_string(r0.bound) + " vs " + std::to_string(r1.bound));
    }
  }
  p.st.push_back(std::move(s));
}

void plan_compile(vk_bplan &p, bool tune) {
  if (p.compiled) usage_error("the plan is already compiled");
  Ctx &c = *p.c;
  const size_t ni = p.nitems;
  p.epoch = static_cast<uint32_t *>(plan_alloc(p, 64));
  fill_u64(c, reinterpret_cast<uint64_t *>(p.epoch), 8, 0);
  for (size_t k = 0; k < p.st.size(); ++k) {
    PlanStmt &s = p.st[k];
    if (s.kind >= 3) {
      // argsort, group_id, reduce_into_add_f64: results at each item's bound (n rows)
      uint64_t total = 0;
      s.off.resize(ni);
      for (size_t i = 0; i < ni; ++i) {
        s.off[i] = total;
        total += s.bound[i];
      }
      s.outbuf = new_array(c, s.d->out_t, total ? total : 1);
      std::vector<BatchItem> items(ni);
      for (size_t i = 0; i < ni; ++i) {
        BatchItem &it = items[i];
        it.a = rows_ref(p, s.args[0], i).ptr;
        it.n = s.bound[i];
        it.out = static_cast<char *>(dptr(s.outbuf)) + s.off[i] * 8;
        if (s.kind == 5) it.b = rows_ref(p, s.args[1], i).ptr;
      }
      if (s.kind == 3) {
        s.sort = plan_sort_new(c, s.d->t0, true, true, items, nullptr, nullptr);
        s.clear.emplace_back(s.sort->status, s.sort->status_words);
      } else if (s.kind == 4) {
        s.group = plan_group_new(c, s.d->t0, items, p.epoch);
        s.counts = s.group->groups;
        s.clear.emplace_back(s.group->sel.status, s.group->status_words);
      } else {
        s.counts = p.st[size_t(s.gsrc)].group->groups;
        s.fold = plan_fold_new(c, items, s.counts);
        s.clear.emplace_back(s.fold->sort->status, s.fold->sort->status_words);
      }
      continue;
    }
    // tile bounds for the smallest tile any candidate uses
    uint64_t unit = 1, per = 1;
    if (s.kind == 0) per = select_tile_rows(8);
    else if (s.kind == 1) unit = 16, per = canon_tile_blocks();
    else per = 4096;
    uint64_t max_tiles = 0;
    for (size_t i = 0; i < ni; ++i) max_tiles += tiles_for(s.bound[i], unit, per);
    if (max_tiles > 0x7FFFFFFFull) usage_error("plan statement too large");
    // device items: pointers and the rows known on the host (device counts replace them)
    std::vector<BatchItem> items(ni);
    if (s.kind == 0) {
      uint64_t total = 0;
      s.off.resize(ni);
      for (size_t i = 0; i < ni; ++i) {
        s.off[i] = total;
        total += s.bound[i];
      }
      s.outbuf = new_array(c, s.d->out_t, total ? total : 1);
      s.counts = static_cast<uint64_t *>(plan_alloc(p, ni * 8));
    }
    const uint32_t w_out = type_width(s.d->out_t);
    for (size_t i = 0; i < ni; ++i) {
      const RowsRef r0 = rows_ref(p, s.args[0], i);
      BatchItem &it = items[i];
      it.a = r0.ptr;
      it.n = r0.src >= 0 ? 0 : r0.bound;
      if ((s.kind == 0 && s.ss.kind == 0) || (s.kind == 1 && s.cs.kind == 1)) it.b = rows_ref(p, s.args[1], i).ptr;
      if (s.kind == 0) it.out = static_cast<char *>(dptr(s.outbuf)) + s.off[i] * w_out;
    }
    s.dp.items = static_cast<BatchItem *>(plan_alloc(p, ni * sizeof(BatchItem)));
    check_hip(hipMemcpyAsync(s.dp.items, items.data(), ni * sizeof(BatchItem), hipMemcpyHostToDevice, c.stream),
              "plan items");
    s.dp.max_tiles = static_cast<uint32_t>(max_tiles);
    s.dp.tile_item = static_cast<uint32_t *>(plan_alloc(p, max_tiles * 4));
    s.dp.ntiles = static_cast<uint32_t *>(plan_alloc(p, 8));
    s.dp.ticket = static_cast<uint32_t *>(plan_alloc(p, 8));
    fill_u64(c, reinterpret_cast<uint64_t *>(s.dp.ticket), 1, 0);
    s.dp.epoch = p.epoch;
    if (s.kind == 0) s.status_words = max_tiles;
    else if (s.kind == 1) s.status_words = max_tiles * canon_status_words_per_tile();
    if (s.status_words) {
      s.dp.status = static_cast<uint64_t *>(plan_alloc(p, s.status_words * 8));
      fill_u64(c, s.dp.status, s.status_words, 0);
      s.clear.emplace_back(s.dp.status, s.status_words);
    }
    if (s.kind == 1) {
      s.res = static_cast<uint64_t *>(plan_alloc(p, ni * 8));
      s.acc = static_cast<double *>(plan_alloc(p, ni * 16 * 8));
      s.zero = static_cast<double *>(plan_alloc(p, ni * 16 * 8));
    } else if (s.kind == 2) {
      s.res = static_cast<uint64_t *>(plan_alloc(p, ni * s.words * 8));
      s.partials = plan_alloc(p, max_tiles * 24);
    }
  }
  check_hip(hipStreamSynchronize(c.stream), "plan setup");
  if (tune) plan_tune(p);
  p.compiled = true;
  if (!kCapturePlans) return;  // then a run enqueues the launch sequence itself
  // capture: kernels only (no allocation, no copy, no memset, no synchronisation)
  check_hip(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    plan_run_eager(p);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  check_hip(hipStreamEndCapture(c.stream, &p.graph), "end capture");
  check_hip(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0), "graph instantiate");
}

void plan_fetch(vk_bplan &p) {
  if (p.fresh) return;
  Ctx &c = *p.c;
  for (auto &s : p.st) {
    if (s.kind == 3) {
      s.host.assign(s.bound.begin(), s.bound.end());
    } else if (s.kind == 0 || s.kind == 4 || s.kind == 5) {
      s.host.resize(p.nitems);
      check_hip(hipMemcpyAsync(s.host.data(), s.counts, p.nitems * 8, hipMemcpyDeviceToHost, c.stream), "counts");
    } else {
      const uint64_t w = s.kind == 1 ? 1 : s.words;
      s.host.resize(p.nitems * w);
      check_hip(hipMemcpyAsync(s.host.data(), s.res, p.nitems * w * 8, hipMemcpyDeviceToHost, c.stream), "results");
    }
  }
  check_hip(hipStreamSynchronize(c.stream), "plan results");
  p.fresh = true;
}

vk_value plan_result(vk_bplan &p, size_t k, size_t i) {
  if (!p.compiled) usage_error("the plan is not compiled");
  if (k >= p.st.size() || i >= p.nitems) usage_error("plan result out of range");
  plan_fetch(p);
  PlanStmt &s = p.st[k];
  vk_value v{VK_VALUE_SCALAR, nullptr, {0, 0}};
  if (s.kind == 0 || s.kind >= 3) {
    // selections and bucket sums: their device row counts; argsort and group_id: n rows
    const uint64_t len = (s.kind == 3 || s.kind == 4) ? s.bound[i] : s.host[i];
    v.kind = VK_VALUE_ARRAY;
    s.outbuf->buf->refs++;
    v.array = new vk_array{1, s.outbuf->buf, s.outbuf->offset + s.off[i] * type_width(s.d->out_t), len, s.d->out_t};
  } else if (s.kind == 1) {
    v.scalar = scalar(VK_F64, s.host[i]);
  } else {
    uint64_t w[3] = {0, 0, 0};
    for (uint32_t q = 0; q < s.words; ++q) w[q] = s.host[i * s.words + q];
    v.scalar = finish_words(*s.d, w);
  }
  return v;
}

void plan_free(vk_bplan *p) {
  if (!p) return;
  Ctx &c = *p->c;
  (void)hipStreamSynchronize(c.stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto &s : p->st) {
    if (s.outbuf) release(s.outbuf);
    plan_sort_free(c, s.sort);
    plan_group_free(c, s.group);
    plan_fold_free(c, s.fold);
  }
  for (void *q : p->allocs) tmp_free(c, q);
  for (vk_array *a : p->inputs) release(a);
  delete p;
  ctx_release(&c);
}

}  // namespace
}  // namespace vk

using namespace vk;

extern "C" {

int vk_ctx_create(int device, vk_ctx **out, vk_status *st) {
  return guarded(st, [&] { *out = create_ctx(device); });
}
void vk_ctx_destroy(vk_ctx *ctx) { destroy_ctx(ctx); }
int vk_ctx_sync(vk_ctx *ctx, vk_status *st) {
  return guarded(st, [&] { check_hip(hipStreamSynchronize(ctx->stream), "sync"); });
}
void *vk_ctx_stream(vk_ctx *ctx) { return ctx->stream; }
void vk_ctx_trim(vk_ctx *ctx) {
  (void)hipStreamSynchronize(ctx->stream);
  ctx->pool->trim();
}
uint64_t vk_ctx_pool_bytes(const vk_ctx *ctx, uint64_t *in_use) {
  if (in_use) *in_use = ctx->pool->in_use();
  return ctx->pool->held();
}
void vk_ctx_set_trace(vk_ctx *ctx, int on) { ctx->trace = on != 0; }
size_t vk_ctx_trace_take(vk_ctx *ctx, char *buf, size_t cap) {
  const size_t n = ctx->trace_buf.size();
  if (buf && cap) {
    const size_t k = n < cap - 1 ? n : cap - 1;
    std::memcpy(buf, ctx->trace_buf.data(), k);
    buf[k] = 0;
  }
  ctx->trace_buf.clear();
  return n;
}
double vk_ctx_last_us(const vk_ctx *ctx) { return ctx->last_us; }

int vk_array_new(vk_ctx *ctx, int32_t type, uint64_t len, vk_array **out, vk_status *st) {
  return guarded(st, [&] {
    if (!valid_type(type)) usage_error("unknown element type");
    *out = new_array(*ctx, type, len);
  });
}
int vk_array_from_host(vk_ctx *ctx, int32_t type, const void *host, uint64_t len, vk_array **out,
                       vk_status *st) {
  return guarded(st, [&] {
    if (!valid_type(type)) usage_error("unknown element type");
    vk_array *a = new_array(*ctx, type, len);
    if (len)
      check_hip(hipMemcpyAsync(dptr(a), host, abytes(a), hipMemcpyHostToDevice, ctx->stream), "upload");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    *out = a;
  });
}
int vk_array_to_host(vk_ctx *ctx, const vk_array *a, void *host, vk_status *st) {
  return guarded(st, [&] {
    if (a->len)
      check_hip(hipMemcpyAsync(host, dptr(a), abytes(a), hipMemcpyDeviceToHost, ctx->stream), "download");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}
int vk_array_wrap(vk_ctx *ctx, int32_t type, void *device_ptr, uint64_t len, vk_array **out,
                  vk_status *st) {
  return guarded(st, [&] {
    if (!valid_type(type)) usage_error("unknown element type");
    if (reinterpret_cast<uintptr_t>(device_ptr) % type_width(type))
      usage_error("device pointer is not aligned for " + tn(type));
    auto *b = new Buffer{ctx, device_ptr, len * type_width(type), false, 1};
    *out = new vk_array{1, b, 0, len, type};
  });
}
int vk_array_slice(const vk_array *a, uint64_t start, uint64_t len, vk_array **out, vk_status *st) {
  return guarded(st, [&] {
    if (start > a->len || len > a->len - start)
      usage_error("slice " + std::to_string(start) + "+" + std::to_string(len) + " out of range for length " +
                  std::to_string(a->len));
    a->buf->refs++;
    *out = new vk_array{1, a->buf, a->offset + start * type_width(a->type), len, a->type};
  });
}
void vk_array_retain(vk_array *a) {
  if (a) a->refs++;
}
void vk_array_release(vk_array *a) { release(a); }
uint64_t vk_array_len(const vk_array *a) { return a->len; }
int32_t vk_array_type(const vk_array *a) { return a->type; }
uint64_t vk_array_bytes(const vk_array *a) { return abytes(a); }
void *vk_array_device_ptr(const vk_array *a) { return dptr(a); }

int32_t vk_kernel_count(void) {
  build_table();
  return static_cast<int32_t>(g_defs.size());
}
int32_t vk_kernel_lookup(const char *name) {
  build_table();
  auto it = g_by_name.find(name ? name : "");
  return it == g_by_name.end() ? -1 : it->second;
}
const char *vk_kernel_name(int32_t id) {
  build_table();
  return id >= 0 && id < int32_t(g_defs.size()) ? g_defs[id].name.c_str() : nullptr;
}
int32_t vk_kernel_class(int32_t id) {
  build_table();
  return id >= 0 && id < int32_t(g_defs.size()) ? g_defs[id].cls : -1;
}
const char *vk_variant_name(void) { return "isa=hip"; }

int vk_call(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, vk_value *out, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    TraceScope ts(*ctx, d.name.c_str(), args, nargs);
    *out = call(*ctx, d, args, nargs);
    ts.done(out->kind == VK_VALUE_ARRAY && out->array ? abytes(out->array) : 0);
  });
}

int vk_reducer_new(vk_ctx *ctx, const char *name, vk_reducer **out, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    if (d.cls != VK_REDUCTION) usage_error(d.name + " is not a reduction");
    *out = reducer_new(*ctx, g_by_name[d.name]);
  });
}
int vk_reducer_update(vk_reducer *r, const vk_arg *args, size_t nargs, vk_status *st) {
  return guarded(st, [&] { reducer_update(r, args, nargs); });
}
int vk_reducer_finish(vk_reducer *r, vk_scalar *out, vk_status *st) {
  return guarded(st, [&] { *out = reducer_finish(r); });
}
void vk_reducer_free(vk_reducer *r) { reducer_free(r); }

int vk_scanner_new(vk_ctx *ctx, const char *name, vk_scanner **out, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    if (d.cls != VK_SCAN) usage_error(d.name + " is not a scan");
    *out = scanner_new(*ctx, g_by_name[d.name]);
  });
}
int vk_scanner_update(vk_scanner *s, const vk_arg *args, size_t nargs, vk_array *out, vk_status *st) {
  return guarded(st, [&] { scanner_update(s, args, nargs, out); });
}
void vk_scanner_free(vk_scanner *s) { scanner_free(s); }

int vk_builder_new(vk_ctx *ctx, int32_t type, vk_builder **out, vk_status *st) {
  return guarded(st, [&] {
    if (!valid_type(type)) usage_error("unknown element type");
    *out = new vk_builder{ctx, type, nullptr, 0, 0};
    ctx_retain(*ctx);
  });
}
uint64_t vk_builder_len(const vk_builder *b) { return b->len; }
int vk_builder_finish(vk_builder *b, vk_array **out, vk_status *st) {
  int rc = guarded(st, [&] { *out = builder_take(b); });
  Ctx *c = b->c;
  if (b->ptr) c->pool->free(b->ptr);  // left only when taking failed
  delete b;
  ctx_release(c);
  return rc;
}
void vk_builder_free(vk_builder *b) {
  if (!b) return;
  Ctx *c = b->c;
  if (b->ptr) c->pool->free(b->ptr);
  delete b;
  ctx_release(c);
}

int vk_select_tile(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, uint64_t base,
                   vk_builder *out, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    if (d.cls != VK_SELECTION) usage_error(d.name + " is not a selection");
    select_tile(*ctx, d, args, nargs, base, out);
  });
}

int vk_region_reduce(vk_ctx *ctx, const vk_rkernel *kernels, size_t nkernels, const char *root,
                     const vk_src *root_args, size_t nroot, const vk_arg *inputs, size_t ninputs, vk_scalar *out,
                     int32_t *fused, vk_status *st) {
  return guarded(st, [&] {
    RegionPlan p;
    const bool one = region_compile(kernels, nkernels, root, root_args, nroot, inputs, ninputs, p);
    if (fused) *fused = one ? 1 : 0;
    // one trace line for a fused pass, named by its root; a kernel-by-kernel run traces each call
    TraceScope ts(*ctx, root ? root : "region", nullptr, 0);
    if (one) {
      ts.n = p.x->len;
      ts.read = abytes(p.x) + (p.y ? abytes(p.y) : 0);
      *out = region_run(*ctx, p);
      ts.done(0);
    } else {
      ts.on = false;
      *out = region_eval(*ctx, kernels, nkernels, root, root_args, nroot, inputs, ninputs);
    }
  });
}

int vk_select_into(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, vk_array *out,
                   uint64_t *kept, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    if (d.cls != VK_SELECTION) usage_error(d.name + " is not a selection");
    const int64_t rows = check_args(d, args, nargs);
    const uint64_t n = rows < 0 ? 0 : static_cast<uint64_t>(rows);
    if (n > VK_SELECT_INTO_MAX_ROWS)
      usage_error("vk_select_into takes at most " + std::to_string(VK_SELECT_INTO_MAX_ROWS) + " rows, got " +
                  std::to_string(n));
    if (!out || out->type != d.out_t)
      usage_error(d.name + " produces " + tn(d.out_t) + " rows, got a " + (out ? tn(out->type) : std::string("null")) +
                  " destination");
    if (out->len < n)
      usage_error("destination of " + std::to_string(out->len) + " rows is shorter than the input's " +
                  std::to_string(n) + " rows");
    if (!kept) usage_error("no place for the kept count");
    Ctx &c = *ctx;
    const IntoScratch sc = select_into_scratch(c, n);
    SelectSpec s{};
    const void *x = nullptr;
    const uint8_t *mask = nullptr;
    if (d.impl == S_COMPRESS) {
      s.kind = 0;
      s.type = d.t0;
      x = dptr(args[0].rows);
      mask = static_cast<const uint8_t *>(dptr(args[1].rows));
    } else if (d.impl == S_FILTER) {
      s.kind = 1;
      s.type = d.t0;
      s.cmp = d.p0;
      s.cbits = args[1].splat.bits;
      x = dptr(args[0].rows);
    } else {
      s.kind = 2;
      s.type = VK_I64;
      mask = static_cast<const uint8_t *>(dptr(args[0].rows));
    }
    select_rows_scratch(c, s, x, mask, n, dptr(out), sc.count, sc.status, sc.epoch, sc.ticket);
    uint64_t got = 0;
    read_back(c, sc.count, 8, &got);
    *kept = got;
  });
}

int vk_elementwise_tile(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, vk_array *out,
                        vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    if (d.cls != VK_ELEMENTWISE) usage_error(d.name + " is not an elementwise kernel");
    if (!out || out->type != d.out_t) usage_error("destination is not aligned for " + tn(d.out_t));
    const int64_t rows = check_args(d, args, nargs);
    if (rows < 0) usage_error("an array kernel needs an array operand");
    if (uint64_t(rows) != out->len)
      usage_error("length mismatch: " + std::to_string(rows) + " rows for a destination of " +
                  std::to_string(out->len) + " rows");
    run_elementwise(*ctx, d, args, nargs, out);
  });
}

void vk_value_release(vk_value *v) {
  if (v && v->kind == VK_VALUE_ARRAY && v->array) {
    release(v->array);
    v->array = nullptr;
  }
}

int vk_batch_call(vk_ctx *ctx, const char *name, const vk_arg *args, size_t nargs, size_t nitems,
                  vk_value *out, vk_status *st) {
  return guarded(st, [&] {
    const Def &d = def_of(name);
    TraceScope ts(*ctx, d.name.c_str(), args, nargs * nitems);
    std::vector<vk_value> v = batch_call(*ctx, d, args, nargs, nitems);
    uint64_t wb = 0;
    for (size_t i = 0; i < nitems; ++i) {
      out[i] = v[i];
      if (v[i].kind == VK_VALUE_ARRAY && v[i].array) wb += abytes(v[i].array);
    }
    ts.done(wb);
  });
}

int vk_bplan_new(vk_ctx *ctx, vk_array *const *inputs, size_t ninputs, size_t nitems, vk_bplan **out,
                 vk_status *st) {
  return guarded(st, [&] {
    if (nitems == 0) usage_error("a plan needs at least one item");
    auto *p = new vk_bplan();
    p->c = ctx;
    p->nitems = nitems;
    p->ninputs = ninputs;
    p->inputs.assign(inputs, inputs + ninputs * nitems);
    for (vk_array *a : p->inputs) {
      if (!a) {
        delete p;
        usage_error("a plan input is null");
      }
    }
    for (vk_array *a : p->inputs) a->refs++;
    ctx_retain(*ctx);
    *out = p;
  });
}

int vk_bplan_add(vk_bplan *plan, const char *name, const vk_bref *args, size_t nargs, size_t *stmt, vk_status *st) {
  return guarded(st, [&] {
    plan_add(*plan, name, args, nargs);
    if (stmt) *stmt = plan->st.size() - 1;
  });
}

int vk_bplan_compile(vk_bplan *plan, int tune, vk_status *st) {
  return guarded(st, [&] { plan_compile(*plan, tune != 0); });
}

int vk_bplan_run(vk_bplan *plan, vk_status *st) {
  return guarded(st, [&] {
    if (!plan->compiled) usage_error("the plan is not compiled");
    if (plan->exec) check_hip(hipGraphLaunch(plan->exec, plan->c->stream), "graph launch");
    else plan_run_eager(*plan);
    plan->fresh = false;
  });
}

int vk_bplan_result(vk_bplan *plan, size_t stmt, size_t item, vk_value *out, vk_status *st) {
  return guarded(st, [&] { *out = plan_result(*plan, stmt, item); });
}

int vk_bplan_counts(vk_bplan *plan, size_t stmt, uint64_t *out, vk_status *st) {
  return guarded(st, [&] {
    if (!plan->compiled) usage_error("the plan is not compiled");
    if (stmt >= plan->st.size()) usage_error("plan result out of range");
    plan_fetch(*plan);
    const PlanStmt &s = plan->st[stmt];
    if (s.kind == 0 || s.kind >= 3) {  // kept rows, rows, group counts, bucket counts
      std::memcpy(out, s.host.data(), plan->nitems * 8);
    } else if (s.kind == 1) {
      std::memcpy(out, s.host.data(), plan->nitems * 8);
    } else {
      for (size_t i = 0; i < plan->nitems; ++i) {
        uint64_t w[3] = {0, 0, 0};
        for (uint32_t q = 0; q < s.words; ++q) w[q] = s.host[i * s.words + q];
        out[i] = finish_words(*s.d, w).bits;
      }
    }
  });
}

size_t vk_bplan_tuning(const vk_bplan *plan, char *buf, size_t cap) {
  std::string all;
  for (const auto &l : plan->tuned) all += l + "\n";
  if (buf && cap) {
    const size_t k = all.size() < cap - 1 ? all.size() : cap - 1;
    std::memcpy(buf, all.data(), k);
    buf[k] = 0;
  }
  return all.size();
}

void vk_bplan_free(vk_bplan *plan) { plan_free(plan); }

}  // extern "C"
