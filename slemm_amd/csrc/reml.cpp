// REML drivers over the device engine (SLDF_REML, LMC_REML of src/reml.cpp:549-882).
#include "common.hpp"

using namespace slemm;

extern "C" int slemm_gpu_sldf_reml(slemm_geno*, const float*, int64_t, int, const float*, int, int, float, float,
                                   float, float, int, int, float*, float*, float*, float*, slemm_reml_stats*) {
  return guard([&] { fail(SLEMM_ERR_STATE, "slemm_gpu_sldf_reml: not available in this build"); });
}

extern "C" int slemm_gpu_lmc_reml(slemm_geno*, const float*, int64_t, int, const float*, int, int, float, float,
                                  float, int, int, float*, float*, float*, slemm_reml_stats*) {
  return guard([&] { fail(SLEMM_ERR_STATE, "slemm_gpu_lmc_reml: not available in this build"); });
}

extern "C" int slemm_gpu_gwas_score(slemm_geno*, const float*, float, float*, float*, void*) {
  return guard([&] { fail(SLEMM_ERR_STATE, "slemm_gpu_gwas_score: not available in this build"); });
}
