// chi2test.cpp -- the chi-square test of BSDF sampling (src/core/Chi2Test.cpp),
// registered as "chi2test".  For every BSDF of the test and `testCount`
// incident directions drawn from a pcg32 stream (its default seed, as the
// reference), `sampleCount` BSDF::Sample directions are drawn on the GPU
// (hk_bsdf_batch) into a contingency table of `resolution` x 2 `resolution`
// cells over (cos theta, phi); the expected frequencies integrate BSDF::Pdf,
// also evaluated on the GPU, over each cell.  Pearson's statistic, with cells
// of low expected frequency pooled, is tested at `significanceLevel`,
// Sidak-corrected over all tests.
//
// Cell integrals: composite Simpson's rule with 128 intervals per axis.  The
// reference integrates adaptively (ext/hypothesis/hypothesis.h:60-101: from 2
// intervals, halving to depth 6, i.e. down to 1/128 of the cell); this uses
// that finest step everywhere, with all evaluations batched in one GPU call
// per group of cells.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "hikari.h"

namespace hikari {

namespace {

// Regularised lower incomplete gamma P(a, x): its power series below a + 1,
// the continued fraction of Q = 1 - P (modified Lentz) above.
double GammaP(double a, double x) {
  if (x <= 0.0) return 0.0;
  double lg = std::lgamma(a);
  if (x < a + 1.0) {
    double term = 1.0 / a, sum = term;
    for (int n = 1; n < 100000; n++) {
      term *= x / (a + n);
      sum += term;
      if (std::fabs(term) < std::fabs(sum) * 1e-16) break;
    }
    return sum * std::exp(-x + a * std::log(x) - lg);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-16) break;
  }
  return 1.0 - std::exp(-x + a * std::log(x) - lg) * h;
}

}  // namespace

double Chi2Cdf(double x, int dof) { return dof > 0 ? GammaP(0.5 * dof, 0.5 * x) : 0.0; }

bool PooledChi2Test(int n_cells, const double *observed, const double *expected, long sample_count, double min_exp,
                    double significance, int n_tests, std::string &msg) {
  // cells in increasing expected frequency; the smallest are pooled until
  // the pool reaches min_exp, the rest contribute one term each
  std::vector<int> order(size_t(std::max(n_cells, 0)));
  for (int i = 0; i < n_cells; i++) order[size_t(i)] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return expected[a] < expected[b]; });
  double pool_obs = 0, pool_exp = 0, chsq = 0;
  int pooled = 0, dof = 0;
  char buf[256];
  for (int i : order) {
    double e = expected[i], o = observed[i];
    if (e < 0) {
      snprintf(buf, sizeof buf, "Reject the null hypothesis: negative expected frequency %g", e);
      msg = buf;
      return false;
    }
    if (e == 0) {
      // a sample in a cell the pdf says is empty: allowed only at the level
      // of rounding (1 in 10^5 of the samples)
      if (o > double(sample_count) * 1e-5) {
        snprintf(buf, sizeof buf, "Reject the null hypothesis: %g samples in a cell with expected frequency 0", o);
        msg = buf;
        return false;
      }
    } else if (e < min_exp || (pool_exp > 0 && pool_exp < min_exp)) {
      pool_obs += o;
      pool_exp += e;
      pooled++;
    } else {
      chsq += (o - e) * (o - e) / e;
      dof++;
    }
  }
  if (pool_exp > 0 || pool_obs > 0) {
    chsq += (pool_obs - pool_exp) * (pool_obs - pool_exp) / pool_exp;
    dof++;
  }
  dof -= 1;  // the frequencies are normalised to the sample count
  if (dof <= 0) {
    snprintf(buf, sizeof buf, "Reject the null hypothesis: %d degrees of freedom", dof);
    msg = buf;
    return false;
  }
  double p = 1.0 - Chi2Cdf(chsq, dof);
  double alpha = 1.0 - std::pow(1.0 - significance, 1.0 / double(std::max(1, n_tests)));
  bool ok = !(p < alpha || !std::isfinite(p));
  snprintf(buf, sizeof buf, "%s: chi^2 = %.10g, %d d.o.f. (%d cells pooled), p-value = %.10g, level %.6g",
           ok ? "Accepted the null hypothesis" : "Reject the null hypothesis", chsq, dof, pooled, p, alpha);
  msg = buf;
  return ok;
}

class Chi2Test : public TestObject {
 public:
  explicit Chi2Test(const PropertyList &p) {
    significance_ = p.GetFloat("significanceLevel", 0.01f);
    res_ = p.GetInteger("resolution", 10);
    min_exp_ = p.GetInteger("minExpFrequency", 5);
    sample_count_ = p.GetInteger("sampleCount", -1);
    test_count_ = p.GetInteger("testCount", 5);
    if (res_ <= 0 || test_count_ < 0) throw HikariException("chi2test: bad resolution or testCount");
    if (sample_count_ < 0) sample_count_ = res_ * 2 * res_ * 5000;  // ~5000 samples per cell
  }
  ~Chi2Test() override {
    for (BSDF *b : bsdfs_) delete b;
  }
  const char *TestKind() const override { return "chi2test"; }
  void AddChild(Object *child, const std::string &name) override {
    if (child->GetClassType() != EClassType::EBSDF)
      throw HikariException("Chi2Test::AddChild(<%s>, <%s>) is not supported!",
                            ClassTypeName(child->GetClassType()).c_str(), name.c_str());
    bsdfs_.push_back(static_cast<BSDF *>(child));
  }
  size_t NumBSDFs() const { return bsdfs_.size(); }
  void Activate() override;
  std::string ToString() const override {
    char b[160];
    snprintf(b, sizeof b, "Chi2Test[significanceLevel = %g, resolution = %d x %d, minExpFrequency = %d, "
             "sampleCount = %d, testCount = %d]", significance_, res_, 2 * res_, min_exp_, sample_count_, test_count_);
    return b;
  }

 private:
  // hk_bsdf_batch over n records: sampled weight and direction for the
  // uniforms u (n x 4), Pdf for the directions wo (n x 3)
  void Batch(const std::vector<float> &records, int idx, size_t n, const float wi[3], const std::vector<float> &wo,
             const std::vector<float> &u, std::vector<float> &pdf, std::vector<float> &f, std::vector<float> &swo) const;

  float significance_;
  int res_, min_exp_, sample_count_, test_count_;
  std::vector<BSDF *> bsdfs_;
};

void Chi2Test::Batch(const std::vector<float> &records, int idx, size_t n, const float wi[3],
                     const std::vector<float> &wo, const std::vector<float> &u, std::vector<float> &pdf,
                     std::vector<float> &f, std::vector<float> &swo) const {
  std::vector<float> wis(n * 3), ev(n * 3), eta(n), spdf(n);
  std::vector<int32_t> meas(n, 1), smeas(n);
  for (size_t k = 0; k < n; k++) std::memcpy(&wis[3 * k], wi, 12);
  pdf.resize(n);
  f.resize(n * 3);
  swo.resize(n * 3);
  CheckHk(hk_bsdf_batch(GpuContext(), records.data(), uint32_t(records.size() / HK_MATERIAL_STRIDE), uint32_t(idx),
                        nullptr, uint32_t(n), wis.data(), wo.data(), meas.data(), u.data(), ev.data(), pdf.data(),
                        f.data(), swo.data(), smeas.data(), eta.data(), spdf.data()),
          "chi2test: BSDF batch");
}

void Chi2Test::Activate() {
  passed = total = 0;
  if (DryRun()) return;
  const int phi_res = 2 * res_, cells = res_ * phi_res;
  const int n_tests = int(bsdfs_.size()) * test_count_;
  constexpr int kIntervals = 128;  // per axis and cell
  constexpr int kPts = kIntervals + 1;
  constexpr int kCellsPerBatch = 16;
  // the reference's stream: per test cos(theta) and phi of the incident
  // direction, then two numbers per sample (src/core/Chi2Test.cpp:84-100),
  // so the incident directions are the reference's; the further numbers a
  // BSDF may draw come from a second stream
  HostPcg32 rng, extra;
  extra.state = 0x5bd1e995u;
  extra.inc = (0x2545f4914f6cdd1dull << 1) | 1u;
  for (BSDF *b : bsdfs_) {
    std::vector<float> records;
    int idx = b->AppendRecords(records);
    for (int l = 0; l < test_count_; l++) {
      total++;
      float cos_t = rng.NextFloat();
      float sin_t = std::sqrt(std::max(0.0f, 1.0f - cos_t * cos_t));
      float ph = 2.0f * float(M_PI) * rng.NextFloat();
      float wi[3] = {std::cos(ph) * sin_t, std::sin(ph) * sin_t, cos_t};
      // observed: sampled directions with a non-zero weight, binned
      size_t n = size_t(sample_count_);
      std::vector<float> u(n * 4), wo(n * 3, 0.0f), pdf, f, swo;
      for (size_t k = 0; k < n; k++) {
        u[4 * k] = rng.NextFloat();
        u[4 * k + 1] = rng.NextFloat();
        u[4 * k + 2] = extra.NextFloat();  // for BSDFs that draw a further number
        u[4 * k + 3] = extra.NextFloat();
      }
      Batch(records, idx, n, wi, wo, u, pdf, f, swo);
      std::vector<double> obs(size_t(cells), 0.0), expv(size_t(cells), 0.0);
      for (size_t k = 0; k < n; k++) {
        if (f[3 * k] == 0.0f && f[3 * k + 1] == 0.0f && f[3 * k + 2] == 0.0f) continue;
        int ct = std::min(std::max(0, int(std::floor((swo[3 * k + 2] * 0.5f + 0.5f) * float(res_)))), res_ - 1);
        float sp = std::atan2(swo[3 * k + 1], swo[3 * k]) * float(0.5 / M_PI);
        if (sp < 0) sp += 1;
        int pb = std::min(std::max(0, int(std::floor(sp * float(phi_res)))), phi_res - 1);
        obs[size_t(ct * phi_res + pb)] += 1;
      }
      // expected: Pdf integrated over each cell (composite Simpson on a
      // kPts x kPts grid), kCellsPerBatch cells per GPU call.  Cell bounds
      // as src/core/Chi2Test.cpp:132-137 computes them (phi in FP32).
      auto cell_bounds = [&](int c, double &ct0, double &ct1, double &p0, double &p1) {
        int i = c / phi_res, j = c % phi_res;
        ct0 = -1.0 + i * 2.0 / res_;
        ct1 = -1.0 + (i + 1) * 2.0 / res_;
        p0 = double(float(j * 2) * float(M_PI) / float(phi_res));
        p1 = double(float((j + 1) * 2) * float(M_PI) / float(phi_res));
      };
      std::vector<double> wgt(kPts);
      for (int k = 0; k < kPts; k++) wgt[size_t(k)] = (k == 0 || k == kIntervals) ? 1.0 : (k % 2 ? 4.0 : 2.0);
      for (int c0 = 0; c0 < cells; c0 += kCellsPerBatch) {
        int c1 = std::min(cells, c0 + kCellsPerBatch);
        size_t m = size_t(c1 - c0) * kPts * kPts;
        std::vector<float> dir(m * 3), uu(m * 4, 0.5f);
        for (int c = c0; c < c1; c++) {
          double ct0, ct1, p0, p1;
          cell_bounds(c, ct0, ct1, p0, p1);
          for (int a = 0; a < kPts; a++) {
            double ctv = ct0 + (ct1 - ct0) * a / kIntervals;
            double st = std::sqrt(1.0f - ctv * ctv);
            for (int q = 0; q < kPts; q++) {
              double phv = p0 + (p1 - p0) * q / kIntervals;
              float *d = &dir[3 * ((size_t(c - c0) * kPts + size_t(a)) * kPts + size_t(q))];
              d[0] = float(st * std::cos(phv));
              d[1] = float(st * std::sin(phv));
              d[2] = float(ctv);
            }
          }
        }
        Batch(records, idx, m, wi, dir, uu, pdf, f, swo);
        for (int c = c0; c < c1; c++) {
          double ct0, ct1, p0, p1;
          cell_bounds(c, ct0, ct1, p0, p1);
          double hx = (ct1 - ct0) / kIntervals, hy = (p1 - p0) / kIntervals;
          double s = 0;
          for (int a = 0; a < kPts; a++)
            for (int q = 0; q < kPts; q++)
              s += wgt[size_t(a)] * wgt[size_t(q)] * double(pdf[(size_t(c - c0) * kPts + size_t(a)) * kPts + size_t(q)]);
          expv[size_t(c)] = s * (hx / 3.0) * (hy / 3.0) * double(n);
        }
      }
      std::string msg;
      if (PooledChi2Test(cells, obs.data(), expv.data(), long(n), double(min_exp_), significance_, n_tests, msg))
        passed++;
      printf("Testing (wi = %.4f, %.4f, %.4f): %s\n  %s\n", wi[0], wi[1], wi[2], b->ToString().c_str(), msg.c_str());
    }
  }
  printf("Passed %d/%d tests.\n", passed, total);
}

REGISTER_CLASS(Chi2Test, "chi2test")

}  // namespace hikari
