// ref_harness.cpp — C entry points over the reference's own header-only
// geometry code, compiled in place from the reference checkout by
// oracle/Makefile (target `ref`) into oracle/_ref/libref.so.  Used only to
// pin the oracle (tests/ and tests/golden/gen_golden.py); nothing ships it.
//
// Wrapped reference code (not copied; included from the checkout):
//   utils::geometry::DCPQuery<double,3>::operator()  util/SpherocylinderCell.h:64-261
//   utils::Quaternion::getDirectionVector            util/Quaternion.cpp:36-38
//   utils::ArrayMath::normalize / cross_product      util/ArrayMath.h:139-142,218-231
#include <array>

#include "ArrayMath.h"
#include "Quaternion.h"
#include "SpherocylinderCell.h"

namespace {
std::array<double, 3> v3(const double* p) { return {p[0], p[1], p[2]}; }
}  // namespace

extern "C" {

void ref_dcp(const double* P0, const double* P1, const double* Q0, const double* Q1, double* out) {
  utils::geometry::DCPSegment3Segment3<double> query;
  auto r = query(v3(P0), v3(P1), v3(Q0), v3(Q1));
  out[0] = r.parameter[0];
  out[1] = r.parameter[1];
  for (int k = 0; k < 3; ++k) {
    out[2 + k] = r.closest[0][k];
    out[5 + k] = r.closest[1][k];
  }
  out[8] = r.sqrDistance;
  out[9] = r.distance;
}

void ref_direction(const double* q, double* t) {
  auto d = utils::Quaternion::getDirectionVector({q[0], q[1], q[2], q[3]});
  t[0] = d[0];
  t[1] = d[1];
  t[2] = d[2];
}

void ref_normalize(const double* a, double* o) {
  auto n = utils::ArrayMath::normalize(v3(a));
  o[0] = n[0];
  o[1] = n[1];
  o[2] = n[2];
}

void ref_cross(const double* a, const double* b, double* o) {
  auto c = utils::ArrayMath::cross_product(v3(a), v3(b));
  o[0] = c[0];
  o[1] = c[1];
  o[2] = c[2];
}

}  // extern "C"
