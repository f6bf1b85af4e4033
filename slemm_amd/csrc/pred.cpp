// --pred: total genetic values of every individual of a PLINK fileset from the SNP effects of an
// estimate file (src/pred.cpp:4-140): fam/bim/estimate parsing and the IID,GEBV output on the
// host, the bed streamed to the device (only SNPs with an effect, up to the last one) and the
// FP64 product on the GPU (slemm_gpu_gebv, reference summation order).
#include <hip/hip_runtime.h>

#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

std::vector<std::string> split_csv(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  std::stringstream ss(line);
  while (std::getline(ss, cur, ',')) out.push_back(cur);
  if (!line.empty() && line.back() == ',') out.emplace_back();
  return out;
}

}  // namespace

using namespace slemm;

extern "C" int slemm_gpu_pred_plink(const char* estimate_file, const char* bfile_prefix, const char* output_file,
                                    int64_t* n_out, int64_t* m_used) {
  return guard([&] {
    REQUIRE(estimate_file && bfile_prefix && output_file, "null argument");
    // marker effects: SNP id in column 0, allele in column 3, effect in column 9 (:18-25)
    std::ifstream ifs(estimate_file);
    if (!ifs.is_open()) fail(SLEMM_ERR_DATA, std::string("Cannot open ") + estimate_file);
    std::string line;
    std::getline(ifs, line);
    std::map<std::string, double> eff;
    std::map<std::string, std::string> allele;
    while (std::getline(ifs, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      const std::vector<std::string> cols = split_csv(line);
      if (cols.size() < 10 || cols[0].empty() || cols[9].empty()) continue;
      try {
        eff[cols[0]] = std::stod(cols[9]);
      } catch (const std::exception&) {
        fail(SLEMM_ERR_DATA, "bad effect value for SNP " + cols[0] + " in " + estimate_file);
      }
      allele[cols[0]] = cols[3];
    }
    ifs.close();
    // individuals: second fam column (:33-45)
    const std::string prefix(bfile_prefix);
    std::ifstream fam(prefix + ".fam");
    if (!fam.is_open()) fail(SLEMM_ERR_DATA, "Cannot open " + prefix + ".fam");
    std::vector<std::string> iid;
    while (std::getline(fam, line)) {
      if (line.size() < 3) continue;
      std::stringstream ss(line);
      std::string fid, id;
      ss >> fid >> id;
      iid.push_back(id);
    }
    // SNPs in bim order: effect and flip when the bim allele 1 differs from the estimate's (:60-82)
    std::ifstream bim(prefix + ".bim");
    if (!bim.is_open()) fail(SLEMM_ERR_DATA, "Cannot open " + prefix + ".bim");
    std::vector<int64_t> snp_idx;
    std::vector<double> alpha;
    std::vector<uint8_t> flip;
    int64_t m_raw = 0;
    std::string chr, snp, cm, bp, a1, a2;
    while (bim >> chr >> snp >> cm >> bp >> a1 >> a2) {
      const auto it = eff.find(snp);
      if (it != eff.end()) {
        snp_idx.push_back(m_raw);
        alpha.push_back(it->second);
        flip.push_back(a1 != allele[snp] ? 1 : 0);
      }
      ++m_raw;
    }
    if (snp_idx.empty()) fail(SLEMM_ERR_DATA, "No SNPs found both in estimate file and in genotype file.");
    const int64_t n = (int64_t)iid.size();
    REQUIRE(n > 0, "empty fam file");
    // genotypes of the SNPs with an effect only; SNPs after the last one are never read (:92-95)
    slemm_geno* h = nullptr;
    const std::string bed = prefix + ".bed";
    if (slemm_gpu_geno_load_bed(bed.c_str(), n, m_raw, nullptr, 0, snp_idx.data(), (int64_t)snp_idx.size(), &h) !=
        SLEMM_OK)
      fail(SLEMM_ERR_DATA, slemm_gpu_last_error());
    struct HGuard {
      slemm_geno* h;
      ~HGuard() { slemm_gpu_geno_destroy(h); }
    } hg{h};
    const int64_t m = (int64_t)snp_idx.size();
    double* dalpha = nullptr;
    double* dgebv = nullptr;
    std::vector<double> tgv(n);
    try {
      HIPCHECK(hipMalloc(&dalpha, sizeof(double) * m));
      HIPCHECK(hipMalloc(&dgebv, sizeof(double) * n));
      HIPCHECK(hipMemcpy(dalpha, alpha.data(), sizeof(double) * m, hipMemcpyHostToDevice));
      if (slemm_gpu_gebv(h, dalpha, flip.data(), dgebv, nullptr) != SLEMM_OK) fail(SLEMM_ERR_HIP, slemm_gpu_last_error());
      HIPCHECK(hipMemcpy(tgv.data(), dgebv, sizeof(double) * n, hipMemcpyDeviceToHost));
    } catch (...) {
      if (dalpha) (void)hipFree(dalpha);
      if (dgebv) (void)hipFree(dgebv);
      throw;
    }
    (void)hipFree(dalpha);
    (void)hipFree(dgebv);
    // IID,GEBV with the default stream formatting (:143-146)
    std::ofstream ofs(output_file);
    if (!ofs.is_open()) fail(SLEMM_ERR_DATA, std::string("Cannot open ") + output_file);
    ofs << "IID,GEBV\n";
    for (int64_t i = 0; i < n; ++i) ofs << iid[i] << ',' << tgv[i] << "\n";
    if (n_out) *n_out = n;
    if (m_used) *m_used = m;
  });
}
