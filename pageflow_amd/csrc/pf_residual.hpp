// The residual program of a join (PF_RES_* in pageflow_join.h): its
// kernel-argument form, the comparison rules, host validation, and the
// per-pair interpreter the partitioned join evaluates inside its count and
// write passes.  The standalone pair filter (join_core.hip) and the
// partitioned join (hash_join.hip) share it.
//
// Comparison semantics follow operators/operator_utils.py:438-478 of the
// reference: integers compare in int64, float64 comparisons in float64 with
// integer columns converted, `=` on a float64 left operand is isclose.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

struct RfProgram {
  const void* col[PF_RES_MAX_COLS];
  int64_t rows[PF_RES_MAX_COLS];
  int32_t type[PF_RES_MAX_COLS];
  int32_t side[PF_RES_MAX_COLS];
  int64_t ins[PF_RES_MAX_INS][3];
  int32_t n_ins;
};

// torch.isclose(a, b, rtol=1e-5, atol=1e-5) in float64: a == b, or a
// finite |a - b| within atol + |rtol * b|, each step rounded once.
__device__ __forceinline__ bool rf_close(double a, double b) {
  if (a == b) return true;
  const double err = fabs(__dsub_rn(a, b));
  const double tol = __dadd_rn(1e-5, fabs(__dmul_rn(1e-5, b)));
  return isfinite(err) && err <= tol;
}

__device__ __forceinline__ bool rf_rel_i(int rel, int64_t a, int64_t b) {
  switch (rel) {
    case PF_REL_EQ: return a == b;
    case PF_REL_NE: return a != b;
    case PF_REL_LT: return a < b;
    case PF_REL_LE: return a <= b;
    case PF_REL_GT: return a > b;
    default: return a >= b;
  }
}

__device__ __forceinline__ bool rf_rel_f(int rel, double a, double b) {
  switch (rel) {
    case PF_REL_EQ: return a == b;
    case PF_REL_NE: return a != b;
    case PF_REL_LT: return a < b;
    case PF_REL_LE: return a <= b;
    case PF_REL_GT: return a > b;
    case PF_REL_GE: return a >= b;
    default: return rf_close(a, b);
  }
}

// A column value widened to 64 bits: integers sign-extended (uint8
// zero-extended), a float64 as its bits.
__device__ __forceinline__ int64_t rf_widen(const void* p, int type, int64_t i) {
  switch (type) {
    case PF_COL_I32: return (int64_t)((const int32_t*)p)[i];
    case PF_COL_I16: return (int64_t)((const int16_t*)p)[i];
    case PF_COL_I8: return (int64_t)((const int8_t*)p)[i];
    case PF_COL_U8: return (int64_t)((const uint8_t*)p)[i];
    default: return ((const int64_t*)p)[i];  // int64, or a float64's bits
  }
}

// Checks the program: opcodes, relations, slots, operand domains and the
// stack depth (at most 32, exactly one result).
inline int rf_validate(const RfProgram& P, int n_cols) {
  int depth = 0;
  for (int k = 0; k < P.n_ins; ++k) {
    const int64_t w0 = P.ins[k][0];
    const int opc = (int)(w0 & 0xff);
    const bool f64 = (w0 & PF_RES_F64) != 0;
    if (opc == PF_RES_CMP) {
      const int rel = (int)((w0 >> 8) & 0xff);
      if (rel < PF_REL_EQ || rel > PF_REL_CLOSE || (rel == PF_REL_CLOSE && !f64))
        return fail("pair_filter: instruction " + std::to_string(k) + " has an unknown relation");
      for (int o = 0; o < 2; ++o) {
        if (!(w0 & (o ? PF_RES_B_COL : PF_RES_A_COL))) continue;
        const int64_t s = P.ins[k][1 + o];
        if (s < 0 || s >= n_cols)
          return fail("pair_filter: instruction " + std::to_string(k) + " names column slot " +
                      std::to_string(s) + " of " + std::to_string(n_cols));
        if (P.type[s] == PF_COL_F64 && !f64)
          return fail("pair_filter: a float64 column compares in float64 only");
      }
      ++depth;
    } else if (opc == PF_RES_TRUE) {
      ++depth;
    } else if (opc == PF_RES_AND || opc == PF_RES_OR) {
      if (depth < 2) return fail("pair_filter: AND/OR on fewer than two results");
      --depth;
    } else if (opc == PF_RES_NOT) {
      if (depth < 1) return fail("pair_filter: NOT on no result");
    } else {
      return fail("pair_filter: instruction " + std::to_string(k) + " has opcode " +
                  std::to_string(opc));
    }
    if (depth > 32) return fail("pair_filter: the program is deeper than 32 results");
  }
  if (depth != 1) return fail("pair_filter: the program leaves " + std::to_string(depth) +
                              " results, not one");
  return 0;
}

// The C ABI's program arrays as an RfProgram, checked (column types, row
// pointers, then the program itself).
inline int rf_program_of(const void* const* cols, const int32_t* col_types,
                         const int32_t* col_sides, const int64_t* col_rows, int32_t n_cols,
                         const int64_t* ins, int32_t n_ins, RfProgram* P) {
  if (n_cols < 0 || n_cols > PF_RES_MAX_COLS)
    return fail("pair_filter: at most " + std::to_string(PF_RES_MAX_COLS) + " columns");
  if (n_ins < 1 || n_ins > PF_RES_MAX_INS)
    return fail("pair_filter: 1 to " + std::to_string(PF_RES_MAX_INS) + " instructions");
  if (!ins || (n_cols > 0 && (!cols || !col_types || !col_sides || !col_rows)))
    return fail("pair_filter: missing program or column arrays");
  *P = RfProgram{};
  for (int s = 0; s < n_cols; ++s) {
    if (col_types[s] < PF_COL_I64 || col_types[s] > PF_COL_F64)
      return fail("pair_filter: column slot " + std::to_string(s) + " has an unknown type");
    if (col_rows[s] < 0 || (col_rows[s] > 0 && !cols[s]))
      return fail("pair_filter: column slot " + std::to_string(s) + " has no rows pointer");
    P->col[s] = cols[s];
    P->rows[s] = col_rows[s];
    P->type[s] = col_types[s];
    P->side[s] = col_sides[s] != 0;
  }
  P->n_ins = n_ins;
  for (int k = 0; k < n_ins; ++k)
    for (int q = 0; q < 3; ++q) P->ins[k][q] = ins[3 * k + q];
  return rf_validate(*P, n_cols);
}

// ------------------------------------------- the join's pair interpreter --
//
// Inside the partitioned join a column slot's value comes from one of:
//   RS_ROW    the column itself, at the pair's build or probe row (any
//             PF_COL_* type);
//   RS_BLDS   the first build column, gathered into partition order and
//             staged in LDS with the build chunk (index: the chunk row);
//   RS_BPART  the second build column in partition order in HBM (index:
//             the chunk's first row + the chunk row);
//   RS_P0/P1  a probe column carried through the probe side's partition
//             (the probe row's value, held with its hash).
// Every source but RS_ROW holds the widened value (rf_widen).
enum : int32_t { RS_ROW = 0, RS_BLDS = 1, RS_BPART = 2, RS_P0 = 3, RS_P1 = 4 };

// The fast form of a program that is one comparison whose column operands
// are staged or carried (RS_BLDS, RS_P0, RS_P1): decoded on the host and
// held in scalar registers.  An operand that is not the staged build column
// is the same for every candidate of a probe row, so it is resolved once a
// row (jr_fast_row: the carried value or the literal, converted to float64
// where the comparison needs it) and a candidate costs one LDS read of the
// build value and one comparison.
struct JrFast {
  int32_t on;       // 1: the program has this form
  int32_t rel, f64;
  int32_t a_bv, b_bv;  // the operand is the staged build column
  int32_t bv_conv;     // ... an integer column compared in float64
  int32_t asel, bsel;  // otherwise: 0 literal, 2 RS_P0, 3 RS_P1
  int32_t aconv, bconv;
  int64_t a, b;        // literals
  // Range form (set by the join for multi-match builds): exactly one
  // operand is the build column and the relation is an order or equality,
  // written "build rrel x" (x: the row operand, side xa = operand a).  A
  // run sorted by the build operand's key (jr_key) keeps a sub-run.
  int32_t range, rrel, xa;
};

// Keys in a comparison's domain whose int64 order is the comparison's
// order: int64 values as they are; float64 values (integer columns
// converted) with -0.0 folded into +0.0 and the sign-magnitude bits made
// order-preserving; NaN, which no order or equality holds for, the largest
// key (kNanKey).
constexpr int64_t kNanKey = INT64_MAX;
__device__ __forceinline__ int64_t jr_key(int64_t v, bool f64, bool conv) {
  if (!f64) return v;
  double d = conv ? (double)v : __longlong_as_double(v);
  if (d != d) return kNanKey;
  if (d == 0.0) d = 0.0;
  uint64_t u = (uint64_t)__double_as_longlong(d);
  u = (u >> 63) ? ~u : (u | (1ULL << 63));
  return (int64_t)(u ^ (1ULL << 63));
}

// "b rel x" on keys; in the float64 domain a NaN key on either side fails
// every relation.
__device__ __forceinline__ bool jr_key_rel(int rel, bool f64, int64_t kb, int64_t kx) {
  if (f64 && (kb == kNanKey || kx == kNanKey)) return false;
  switch (rel) {
    case PF_REL_EQ: return kb == kx;
    case PF_REL_LT: return kb < kx;
    case PF_REL_LE: return kb <= kx;
    case PF_REL_GT: return kb > kx;
    default: return kb >= kx;
  }
}

struct JoinResidual {
  RfProgram P;
  int32_t src[PF_RES_MAX_COLS];
  const int64_t* bpart1;  // the RS_BPART column in partition order
  JrFast fast;
};

// The fast form of R's program (on = 0 when it has none).
inline JrFast jr_fast_of(const JoinResidual& R) {
  JrFast f{};
  if (R.P.n_ins != 1) return f;
  const int64_t w0 = R.P.ins[0][0];
  if ((w0 & 0xff) != PF_RES_CMP) return f;
  f.rel = (int32_t)((w0 >> 8) & 0xff);
  f.f64 = (w0 & PF_RES_F64) != 0;
  int32_t* bv[2] = {&f.a_bv, &f.b_bv};
  int32_t* sel[2] = {&f.asel, &f.bsel};
  int32_t* conv[2] = {&f.aconv, &f.bconv};
  for (int o = 0; o < 2; ++o) {
    if (!(w0 & (o ? PF_RES_B_COL : PF_RES_A_COL))) continue;
    const int s = (int)R.P.ins[0][1 + o];
    const bool cv = f.f64 && R.P.type[s] != PF_COL_F64;
    switch (R.src[s]) {
      case RS_BLDS: *bv[o] = 1; f.bv_conv = cv; break;
      case RS_P0: *sel[o] = 2; *conv[o] = cv; break;
      case RS_P1: *sel[o] = 3; *conv[o] = cv; break;
      default: return JrFast{};
    }
  }
  f.a = R.P.ins[0][1];
  f.b = R.P.ins[0][2];
  f.on = 1;
  // the range form: one build operand, an order or equality (isclose and
  // <> keep no single sub-run); x rel' b is written b rel x
  if ((f.a_bv != 0) != (f.b_bv != 0) && f.rel != PF_REL_NE && f.rel != PF_REL_CLOSE) {
    static const int32_t flip[6] = {PF_REL_EQ, PF_REL_NE, PF_REL_GT, PF_REL_GE, PF_REL_LT,
                                    PF_REL_LE};
    f.range = 1;
    f.xa = f.b_bv;  // the row operand is a when the build column is b
    f.rrel = f.a_bv ? f.rel : flip[f.rel];
  }
  return f;
}

// The operands of one pair.
struct PairCtx {
  const int64_t* bv;    // RS_BLDS values of the chunk (LDS)
  const int32_t* rows;  // the chunk's build row ids (LDS)
  int li;               // the build row's index in its chunk
  int64_t cb;           // the chunk's first row in partition order
  int64_t prow;         // the probe row
  int64_t pv0, pv1;     // its carried values
};

__device__ __forceinline__ int64_t jr_operand(const JoinResidual& R, int64_t word, bool is_col,
                                              bool f64, const PairCtx& c) {
  if (!is_col) return word;
  const int s = (int)word;
  int64_t v;
  switch (R.src[s]) {
    case RS_BLDS: v = c.bv[c.li]; break;
    case RS_BPART: v = R.bpart1[c.cb + c.li]; break;
    case RS_P0: v = c.pv0; break;
    case RS_P1: v = c.pv1; break;
    default:
      v = rf_widen(R.P.col[s], R.P.type[s], R.P.side[s] ? c.prow : (int64_t)c.rows[c.li]);
      break;
  }
  if (f64 && R.P.type[s] != PF_COL_F64) v = __double_as_longlong((double)v);
  return v;
}

// The program's result on one pair (the instruction words are uniform
// across the wave: kernel arguments).
__device__ __forceinline__ bool jr_eval(const JoinResidual& R, const PairCtx& c) {
  uint32_t stk = 0;
  for (int pc = 0; pc < R.P.n_ins; ++pc) {
    const int64_t w0 = R.P.ins[pc][0];
    const int opc = (int)(w0 & 0xff);
    if (opc == PF_RES_CMP) {
      const int rel = (int)((w0 >> 8) & 0xff);
      const bool f64 = (w0 & PF_RES_F64) != 0;
      const int64_t a = jr_operand(R, R.P.ins[pc][1], (w0 & PF_RES_A_COL) != 0, f64, c);
      const int64_t b = jr_operand(R, R.P.ins[pc][2], (w0 & PF_RES_B_COL) != 0, f64, c);
      const bool r = f64 ? rf_rel_f(rel, __longlong_as_double(a), __longlong_as_double(b))
                         : rf_rel_i(rel, a, b);
      stk = (stk << 1) | (uint32_t)r;
    } else if (opc == PF_RES_AND) {
      const uint32_t t = stk & 1u;
      stk >>= 1;
      stk &= ~1u | t;
    } else if (opc == PF_RES_OR) {
      const uint32_t t = stk & 1u;
      stk = (stk >> 1) | t;
    } else if (opc == PF_RES_NOT) {
      stk ^= 1u;
    } else {  // PF_RES_TRUE
      stk = (stk << 1) | 1u;
    }
  }
  return (stk & 1u) != 0;
}

// The whole program decoded once per workgroup into LDS (1.5 KB): per
// instruction one packed word (opcode, relation, float64 flag, each
// operand's source and whether an integer column converts to float64), the
// operands' column slots, and the literals.  The join's evaluator then
// walks it with uniform LDS reads and no dependent chain of kernel-argument
// loads, and its code appears once per kernel (a loop, not an unrolled
// interpreter), so the kernels stay small enough for the instruction cache.
struct JrLds {
  uint32_t w[PF_RES_MAX_INS];  // opc | rel << 8 | f64 << 16 | asrc << 17 | bsrc << 21
                               //  | aconv << 25 | bconv << 26   (src: 0 literal, RS_* + 1)
  uint32_t slots[PF_RES_MAX_INS];  // aslot | bslot << 8
  int64_t lit[PF_RES_MAX_INS][2];
};

// Thread pc of the block decodes instruction pc; the caller synchronises.
__device__ __forceinline__ void jr_decode(const JoinResidual& R, JrLds* d) {
  const int pc = threadIdx.x;
  if (pc >= R.P.n_ins) return;
  const int64_t w0 = R.P.ins[pc][0], aw = R.P.ins[pc][1], bw = R.P.ins[pc][2];
  const bool ac = (w0 & PF_RES_A_COL) != 0, bc = (w0 & PF_RES_B_COL) != 0;
  const bool f64 = (w0 & PF_RES_F64) != 0;
  const int as = ac ? (int)aw : 0, bs = bc ? (int)bw : 0;
  const uint32_t asrc = ac ? (uint32_t)R.src[as] + 1u : 0u;
  const uint32_t bsrc = bc ? (uint32_t)R.src[bs] + 1u : 0u;
  const bool aconv = ac && f64 && R.P.type[as] != PF_COL_F64;
  const bool bconv = bc && f64 && R.P.type[bs] != PF_COL_F64;
  d->w[pc] = (uint32_t)(w0 & 0xffff) | (uint32_t)f64 << 16 | asrc << 17 | bsrc << 21 |
             (uint32_t)aconv << 25 | (uint32_t)bconv << 26;
  d->slots[pc] = (uint32_t)as | (uint32_t)bs << 8;
  d->lit[pc][0] = aw;
  d->lit[pc][1] = bw;
}

// An operand: src 0 is the literal, else RS_* + 1.
__device__ __forceinline__ int64_t jr_fetch(const JoinResidual& R, uint32_t src, int slot,
                                            bool conv, int64_t lit, const PairCtx& c) {
  int64_t v;
  switch (src) {
    case 0: return lit;
    case RS_BLDS + 1: v = c.bv[c.li]; break;
    case RS_BPART + 1: v = R.bpart1[c.cb + c.li]; break;
    case RS_P0 + 1: v = c.pv0; break;
    case RS_P1 + 1: v = c.pv1; break;
    default:
      v = rf_widen(R.P.col[slot], R.P.type[slot],
                   R.P.side[slot] ? c.prow : (int64_t)c.rows[c.li]);
      break;
  }
  return conv ? __double_as_longlong((double)v) : v;
}

// A comparison without a switch: the orderings (lt, eq, gt) against the
// relation's mask (nibble rel of 0x316482: EQ 2, LT 4, LE 6, GT 1, GE 3;
// NE, 8, is "not equal", so NaN != x holds as in torch); isclose apart.
__device__ __forceinline__ bool jr_compare(int rel, bool f64, int64_t a, int64_t b) {
  bool lt, eq, gt;
  if (f64) {
    const double x = __longlong_as_double(a), y = __longlong_as_double(b);
    if (rel == PF_REL_CLOSE) return rf_close(x, y);
    lt = x < y;
    eq = x == y;
    gt = x > y;
  } else {
    lt = a < b;
    eq = a == b;
    gt = a > b;
  }
  const uint32_t m = (0x316482u >> (4 * rel)) & 0xfu;
  const uint32_t o = (uint32_t)lt << 2 | (uint32_t)eq << 1 | (uint32_t)gt;
  return (m & 8u) ? !eq : (m & o) != 0u;
}

// A fast form's operand that does not read the build column: the literal
// or a carried value of the probe row, as the comparison reads it.
__device__ __forceinline__ int64_t jr_fast_row(int32_t sel, int32_t conv, int64_t lit,
                                               int64_t pv0, int64_t pv1) {
  const int64_t v = sel == 2 ? pv0 : sel == 3 ? pv1 : lit;
  return conv ? __double_as_longlong((double)v) : v;
}

// The fast form on one candidate: bvl the staged build value, ra / rb the
// row's resolved operands.
__device__ __forceinline__ bool jr_fast_eval(const JrFast& f, int64_t bvl, int64_t ra,
                                             int64_t rb) {
  if (f.bv_conv) bvl = __double_as_longlong((double)bvl);
  const int64_t a = f.a_bv ? bvl : ra;
  const int64_t b = f.b_bv ? bvl : rb;
  return jr_compare(f.rel, f.f64 != 0, a, b);
}

// The decoded program's result on one pair.
__device__ __forceinline__ bool jr_eval_lds(const JoinResidual& R, const JrLds* d,
                                            const PairCtx& c) {
  uint32_t stk = 0;
  const int n = R.P.n_ins;
#pragma unroll 1
  for (int pc = 0; pc < n; ++pc) {
    const uint32_t w = d->w[pc];
    const int opc = (int)(w & 0xff);
    if (opc == PF_RES_CMP) {
      const uint32_t sl = d->slots[pc];
      const int64_t a = jr_fetch(R, (w >> 17) & 0xf, (int)(sl & 0xff), ((w >> 25) & 1u) != 0,
                                 d->lit[pc][0], c);
      const int64_t b = jr_fetch(R, (w >> 21) & 0xf, (int)(sl >> 8), ((w >> 26) & 1u) != 0,
                                 d->lit[pc][1], c);
      stk = (stk << 1) | (uint32_t)jr_compare((int)((w >> 8) & 0xff), ((w >> 16) & 1u) != 0, a, b);
    } else if (opc == PF_RES_AND) {
      const uint32_t t = stk & 1u;
      stk >>= 1;
      stk &= ~1u | t;
    } else if (opc == PF_RES_OR) {
      const uint32_t t = stk & 1u;
      stk = (stk >> 1) | t;
    } else if (opc == PF_RES_NOT) {
      stk ^= 1u;
    } else {  // PF_RES_TRUE
      stk = (stk << 1) | 1u;
    }
  }
  return (stk & 1u) != 0;
}

}  // namespace pf
