// mpcd_rev.h -- reverse-mode arithmetic for the adjoint of the solvent step.
//
// An Rv carries a value and the index of the tape entry that made it (-1: a
// constant).  The kernels of the step are templates on their arithmetic type;
// instantiated with Rv, a lane computes the values the FP64 code computes,
// bit for bit (the same operations in the same order), and records on its own
// tape every operation on non-constant operands with the operation's partial
// derivatives.  Comparisons and integer conversions look at the value only,
// so every discrete decision of the step (cell assignment, wall events,
// Metropolis acceptances, activity side, eigen branches) is the forward run's:
// the reverse sweep gives the pathwise derivative.
//
// What a lane reads from a state array through TA<Rv>::ld becomes a leaf
// whose adjoint is added (atomically) to the adjoint array at the same index;
// what it stores through TA<Rv>::st is an output, seeded from the adjoint
// array at that index, which the sweep then clears (an in-place update reads
// the adjoint of its output and leaves the adjoint of its input).  A load
// after a store to the same element returns the stored value's entry.  Lanes
// of one launch write only their own cells' or particles' elements, so the
// clear and the deposits of different lanes never touch the same element.
//
// The tape of one launch is interleaved: entry k of lane l lives at
// k * stride + l.  A counting launch (tape.counting = 1) runs the same code,
// writes nothing and reports the largest number of entries and outputs of any
// lane, so that the recording launch can be given a tape that fits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpcd_dual.h"

namespace mpcd_ad {

constexpr int kRevBlock = 256;  // lanes per workgroup of every reverse launch

struct RevTape {
  int2 *par;      // parents (a, b) of entry k; a = -2 marks a leaf
  double2 *dp;    // d(entry)/d(a), d(entry)/d(b); a leaf keeps its adjoint's address in x
  double *adj;    // adjoints of the entries (the sweep)
  double **optr;  // outputs: the adjoint element of a stored value ...
  int *ovar;      // ... and the entry stored there (-1: a constant)
  int cap, ocap, stride;
  int counting;
  int *need;  // [3]: most entries, most outputs of a lane; [2] != 0: a lane exceeded the tape
};

__device__ __forceinline__ RevTape &rev_tape() {
  __shared__ RevTape s;
  return s;
}
// entries and outputs of this lane so far
__device__ __forceinline__ int2 &rev_cnt() {
  __shared__ int2 s[kRevBlock];
  return s[threadIdx.x];
}
__device__ __forceinline__ size_t rev_lane() { return (size_t)blockIdx.x * blockDim.x + threadIdx.x; }

struct Rv {
  double v;
  int i;
  __device__ __forceinline__ Rv() : v(0.0), i(-1) {}
  __device__ __forceinline__ Rv(double x) : v(x), i(-1) {}  // NOLINT: constants promote implicitly
  __device__ __forceinline__ Rv(double x, int k) : v(x), i(k) {}
  __device__ __forceinline__ Rv &operator+=(const Rv &b);
  __device__ __forceinline__ Rv &operator-=(const Rv &b);
  __device__ __forceinline__ Rv &operator*=(const Rv &b);
  __device__ __forceinline__ Rv &operator/=(const Rv &b);
};

__device__ __forceinline__ int rv_entry(int a, double da, int b, double db) {
  int2 &c = rev_cnt();
  const int k = c.x++;
  const RevTape &T = rev_tape();
  if (!T.counting && k < T.cap) {
    const size_t s = (size_t)k * T.stride + rev_lane();
    T.par[s] = make_int2(a, b);
    T.dp[s] = make_double2(da, db);
  }
  return k;
}

__device__ __forceinline__ Rv rv_un(double v, const Rv &a, double da) {
  return a.i < 0 ? Rv(v) : Rv(v, rv_entry(a.i, da, -1, 0.0));
}
__device__ __forceinline__ Rv rv_bin(double v, const Rv &a, double da, const Rv &b, double db) {
  if (a.i < 0) return b.i < 0 ? Rv(v) : Rv(v, rv_entry(b.i, db, -1, 0.0));
  if (b.i < 0) return Rv(v, rv_entry(a.i, da, -1, 0.0));
  return Rv(v, rv_entry(a.i, da, b.i, db));
}

// a value whose adjoint goes to *adj (null: a constant)
__device__ __forceinline__ Rv rv_leaf(double v, double *adj) {
  if (!adj) return Rv(v);
  return Rv(v, rv_entry(-2, __builtin_bit_cast(double, (uint64_t)(uintptr_t)adj), -2, 0.0));
}

// the entry last stored to the adjoint element *adj by this lane
__device__ __forceinline__ bool rv_stored(const double *adj, int &var) {
  const RevTape &T = rev_tape();
  if (T.counting) return false;
  const int n = min(rev_cnt().y, T.ocap);
  const size_t l = rev_lane();
  for (int o = 0; o < n; o++)
    if (T.optr[(size_t)o * T.stride + l] == adj) {
      var = T.ovar[(size_t)o * T.stride + l];
      return true;
    }
  return false;
}

__device__ __forceinline__ void rv_store(double *adj, const Rv &x) {
  if (!adj) return;
  const RevTape &T = rev_tape();
  int2 &c = rev_cnt();
  if (T.counting) {
    c.y++;
    return;
  }
  const size_t l = rev_lane();
  const int n = min(c.y, T.ocap);
  for (int o = 0; o < n; o++)
    if (T.optr[(size_t)o * T.stride + l] == adj) {
      T.ovar[(size_t)o * T.stride + l] = x.i;
      return;
    }
  const int o = c.y++;
  if (o < T.ocap) {
    T.optr[(size_t)o * T.stride + l] = adj;
    T.ovar[(size_t)o * T.stride + l] = x.i;
  }
}

// every lane of the workgroup: start a tape (all lanes must reach it)
__device__ __forceinline__ void rev_begin(const RevTape &t) {
  if (threadIdx.x == 0) rev_tape() = t;
  rev_cnt() = make_int2(0, 0);
  __syncthreads();
}

// this lane's reverse sweep: seed the outputs from their adjoint elements,
// clear those elements, propagate, and add the leaves' adjoints to theirs
__device__ __forceinline__ void rev_finish() {
  const RevTape &T = rev_tape();
  const int2 c = rev_cnt();
  if (T.counting || c.x > T.cap || c.y > T.ocap) {
    atomicMax(&T.need[0], c.x);
    atomicMax(&T.need[1], c.y);
    if (!T.counting) atomicOr(&T.need[2], 1);
    return;
  }
  const size_t l = rev_lane(), S = T.stride;
  for (int k = 0; k < c.x; k++) T.adj[(size_t)k * S + l] = 0.0;
  for (int o = 0; o < c.y; o++) {
    const int v = T.ovar[(size_t)o * S + l];
    if (v >= 0) T.adj[(size_t)v * S + l] += *T.optr[(size_t)o * S + l];
  }
  for (int o = 0; o < c.y; o++) *T.optr[(size_t)o * S + l] = 0.0;
  for (int k = c.x - 1; k >= 0; k--) {
    const double a = T.adj[(size_t)k * S + l];
    if (a == 0.0) continue;
    const int2 p = T.par[(size_t)k * S + l];
    const double2 d = T.dp[(size_t)k * S + l];
    if (p.x == -2) {
      atomicAdd((double *)(uintptr_t)__builtin_bit_cast(uint64_t, d.x), a);
    } else {
      if (p.x >= 0) T.adj[(size_t)p.x * S + l] += a * d.x;
      if (p.y >= 0) T.adj[(size_t)p.y * S + l] += a * d.y;
    }
  }
}

__device__ __forceinline__ double val(const Rv &a) { return a.v; }

__device__ __forceinline__ Rv operator-(const Rv &a) { return rv_un(-a.v, a, -1.0); }
__device__ __forceinline__ Rv operator+(const Rv &a, const Rv &b) { return rv_bin(a.v + b.v, a, 1.0, b, 1.0); }
__device__ __forceinline__ Rv operator-(const Rv &a, const Rv &b) { return rv_bin(a.v - b.v, a, 1.0, b, -1.0); }
__device__ __forceinline__ Rv operator*(const Rv &a, const Rv &b) { return rv_bin(a.v * b.v, a, b.v, b, a.v); }
__device__ __forceinline__ Rv operator/(const Rv &a, const Rv &b) {
  const double q = a.v / b.v;
  return rv_bin(q, a, 1.0 / b.v, b, -q / b.v);
}
__device__ __forceinline__ Rv operator+(const Rv &a, double b) { return rv_un(a.v + b, a, 1.0); }
__device__ __forceinline__ Rv operator+(double a, const Rv &b) { return rv_un(a + b.v, b, 1.0); }
__device__ __forceinline__ Rv operator-(const Rv &a, double b) { return rv_un(a.v - b, a, 1.0); }
__device__ __forceinline__ Rv operator-(double a, const Rv &b) { return rv_un(a - b.v, b, -1.0); }
__device__ __forceinline__ Rv operator*(const Rv &a, double b) { return rv_un(a.v * b, a, b); }
__device__ __forceinline__ Rv operator*(double a, const Rv &b) { return rv_un(a * b.v, b, a); }
__device__ __forceinline__ Rv operator/(const Rv &a, double b) { return rv_un(a.v / b, a, 1.0 / b); }
__device__ __forceinline__ Rv operator/(double a, const Rv &b) {
  const double q = a / b.v;
  return rv_un(q, b, -q / b.v);
}
__device__ __forceinline__ Rv &Rv::operator+=(const Rv &b) { return *this = *this + b; }
__device__ __forceinline__ Rv &Rv::operator-=(const Rv &b) { return *this = *this - b; }
__device__ __forceinline__ Rv &Rv::operator*=(const Rv &b) { return *this = *this * b; }
__device__ __forceinline__ Rv &Rv::operator/=(const Rv &b) { return *this = *this / b; }

#define MPCD_RV_CMP(op)                                                                           \
  __device__ __forceinline__ bool operator op(const Rv &a, const Rv &b) { return a.v op b.v; } \
  __device__ __forceinline__ bool operator op(const Rv &a, double b) { return a.v op b; }      \
  __device__ __forceinline__ bool operator op(double a, const Rv &b) { return a op b.v; }
MPCD_RV_CMP(<)
MPCD_RV_CMP(>)
MPCD_RV_CMP(<=)
MPCD_RV_CMP(>=)
MPCD_RV_CMP(==)
MPCD_RV_CMP(!=)
#undef MPCD_RV_CMP

// the reference's tolerance tests (mtools.c:61, 74) on values
__device__ __forceinline__ bool feq(const Rv &x, const Rv &y) { return ::fabs(x.v - y.v) <= 1e-8; }
__device__ __forceinline__ bool feq(const Rv &x, double y) { return ::fabs(x.v - y) <= 1e-8; }
__device__ __forceinline__ bool fneq(const Rv &x, const Rv &y) { return ::fabs(x.v - y.v) >= 1e-8; }
__device__ __forceinline__ bool fneq(const Rv &x, double y) { return ::fabs(x.v - y) >= 1e-8; }

// elementary functions: the value by the library routine, the partial
// derivative by its formula (the same ones as the tangent's)
__device__ __forceinline__ Rv sqrt(const Rv &a) {
  const double s = ::sqrt(a.v);
  return rv_un(s, a, s > 0.0 ? 0.5 / s : 0.0);
}
__device__ __forceinline__ Rv sin(const Rv &a) { return rv_un(::sin(a.v), a, ::cos(a.v)); }
__device__ __forceinline__ Rv cos(const Rv &a) { return rv_un(::cos(a.v), a, -::sin(a.v)); }
__device__ __forceinline__ Rv exp(const Rv &a) {
  const double e = ::exp(a.v);
  return rv_un(e, a, e);
}
__device__ __forceinline__ Rv acos(const Rv &a) {
  const double s = 1.0 - a.v * a.v;
  return rv_un(::acos(a.v), a, s > 0.0 ? -1.0 / ::sqrt(s) : 0.0);
}
__device__ __forceinline__ Rv atan(const Rv &a) { return rv_un(::atan(a.v), a, 1.0 / (1.0 + a.v * a.v)); }
__device__ __forceinline__ Rv tanh(const Rv &a) {
  const double t = ::tanh(a.v);
  return rv_un(t, a, 1.0 - t * t);
}
__device__ __forceinline__ Rv pow(const Rv &a, double p) {
  const double r = ::pow(a.v, p);
  return rv_un(r, a, a.v != 0.0 ? p * r / a.v : 0.0);
}
__device__ __forceinline__ Rv fabs(const Rv &a) { return a.v < 0.0 ? -a : a; }

}  // namespace mpcd_ad
