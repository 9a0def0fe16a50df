// ORACLE — test infrastructure only. Entry points of the op-counting build
// (liborc_count.so): serial stepping that returns the FP64 operations spent.
#include <cstring>

#include "../include/pippin_hip.h"
#include "engine.hpp"

namespace orc {
thread_local unsigned long long g_ops = 0;
}
using namespace orc;

struct orcc_sim {
  Model m;
  std::vector<Data> envs;
};

extern "C" {

orcc_sim* orcc_create(const ph_model_desc* d, int64_t n) {
  orcc_sim* s = new orcc_sim();
  s->m = model_from_desc(*d);
  Data proto = data_new(s->m);
  kinematics(s->m, proto);
  s->envs.assign((size_t)n, proto);
  return s;
}

void orcc_destroy(orcc_sim* s) { delete s; }

int orcc_set(orcc_sim* s, int field, int64_t begin, int64_t end, const double* src) {
  for (int64_t i = begin; i < end; i++) {
    Data& d = s->envs[(size_t)i];
    std::vector<Real>& v = field == PH_FIELD_QPOS ? d.qpos : (field == PH_FIELD_QVEL ? d.qvel : d.ctrl);
    for (size_t k = 0; k < v.size(); k++) v[k] = src[(i - begin) * v.size() + k];
  }
  return 0;
}

// Step envs [begin, end) nstep times (step-synchronous, then kinematics as
// PippinCpu::step does) and return the FP64 operations of the steps alone.
int orcc_step(orcc_sim* s, int64_t begin, int64_t end, int64_t nstep, unsigned long long* ops) {
  g_ops = 0;
  for (int64_t t = 0; t < nstep; t++)
    for (int64_t i = begin; i < end; i++)
      if (!step(s->m, s->envs[(size_t)i])) return PH_ERR_NOTPD;
  *ops = g_ops;
  for (int64_t i = begin; i < end; i++) kinematics(s->m, s->envs[(size_t)i]);
  return 0;
}

}  // extern "C"
