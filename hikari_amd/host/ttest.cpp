// ttest.cpp -- the Student's t-test harness (src/core/StudentTTest.cpp),
// registered as "ttest".  Scene tests draw `sampleCount` path samples of the
// scene's (1 x 1) film with the GPU integrator and compare the mean
// luminance with each reference; BSDF tests draw `sampleCount` BSDF::Sample
// weights per incidence angle on the GPU (hk_bsdf_batch).  The test is
// two-sided at `significanceLevel`, Sidak-corrected over the references.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "hikari.h"

namespace hikari {

bool RenderPixelSamples(Integrator *integ, const Scene *scene, uint32_t seed, uint32_t n, std::vector<float> &out);

namespace {

// Regularised incomplete beta function I_x(a, b) by its continued fraction
// (modified Lentz evaluation), with the symmetry I_x(a,b) = 1 - I_{1-x}(b,a)
// to stay in the fast-converging region.
double betacf(double a, double b, double x) {
  const double tiny = 1e-300, eps = 1e-15;
  double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (std::fabs(d) < tiny) d = tiny;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 100000; m++) {
    int m2 = 2 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < eps) break;
  }
  return h;
}

double inc_beta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  double lbt = std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log1p(-x);
  double bt = std::exp(lbt);
  if (x < (a + 1.0) / (a + b + 2.0)) return bt * betacf(a, b, x) / a;
  return 1.0 - bt * betacf(b, a, 1.0 - x) / b;
}

}  // namespace

bool StudentsTTest(double mean, double variance, double reference, long n, double significance, int n_tests,
                   std::string &msg) {
  // The statistic floors the sample variance at 1e-5 (ext/hypothesis/hypothesis.h:324),
  // so a deterministic estimator (zero variance) whose mean matches the reference up to
  // float rounding is accepted rather than given an infinite t.
  double nu = double(n - 1);
  double t = std::fabs(mean - reference) * std::sqrt(double(n) / std::max(variance, 1e-5));
  // two-sided p-value of Student's t with nu degrees of freedom
  double p = std::isinf(t) ? 0.0 : inc_beta(0.5 * nu, 0.5, nu / (nu + t * t));
  double alpha = 1.0 - std::pow(1.0 - significance, 1.0 / double(std::max(1, n_tests)));
  bool ok = !(p < alpha);  // rejected only when p < alpha (hypothesis.h:342)
  char buf[256];
  snprintf(buf, sizeof buf, "%s: sample mean %.6g vs reference %.6g (t = %.6g, p-value = %.6g, level %.6g)",
           ok ? "Accepted the null hypothesis" : "Reject the null hypothesis", mean, reference, t, p, alpha);
  msg = buf;
  return ok;
}

StudentTTest::StudentTTest(const PropertyList &p) : TestObject() {
  significance_ = p.GetFloat("significanceLevel", 0.01f);
  std::string angles = p.GetString("angles", ""), refs = p.GetString("references", "");
  auto parse = [](const std::string &s, std::vector<float> &out) {
    std::string cur;
    for (char ch : s + ",") {
      if (ch == ',' || ch == ' ' || ch == '\t' || ch == '\n') {
        if (!cur.empty()) out.push_back(std::stof(cur));
        cur.clear();
      } else {
        cur += ch;
      }
    }
  };
  parse(angles, angles_);
  parse(refs, references_);
  sample_count_ = p.GetInteger("sampleCount", 100000);
}

StudentTTest::~StudentTTest() {
  for (BSDF *b : bsdfs_) delete b;
  for (Scene *s : scenes_) delete s;
}

void StudentTTest::AddChild(Object *child, const std::string &name) {
  switch (child->GetClassType()) {
    case EClassType::EBSDF: bsdfs_.push_back(static_cast<BSDF *>(child)); break;
    case EClassType::EScene: scenes_.push_back(static_cast<Scene *>(child)); break;
    default:
      throw HikariException("StudentTTest::AddChild(<%s>, <%s>) is not supported!",
                            ClassTypeName(child->GetClassType()).c_str(), name.c_str());
  }
}

static bool g_dry_run = false;
void SetDryRun(bool dry) { g_dry_run = dry; }
bool DryRun() { return g_dry_run; }

void StudentTTest::Activate() {
  passed = total = 0;
  if (g_dry_run) return;
  if (!bsdfs_.empty()) {
    if (references_.size() * bsdfs_.size() != angles_.size() && references_.size() != angles_.size() * bsdfs_.size())
      throw HikariException("Specified a different number of angles and reference values!");
    if (!scenes_.empty()) throw HikariException("Cannot test BSDFs and scenes at the same time!");
    size_t ref = 0;
    // the reference's stream: two numbers per sample from a default-seeded
    // pcg32 (src/core/StudentTTest.cpp:105), so the harness draws the
    // reference's sample points; the further numbers a BSDF may draw come
    // from a second stream
    HostPcg32 rng, extra;
    extra.state = 0x5bd1e995u;
    extra.inc = (0x2545f4914f6cdd1dull << 1) | 1u;
    for (BSDF *b : bsdfs_) {
      std::vector<float> records;
      int idx = b->AppendRecords(records);
      for (size_t i = 0; i < angles_.size() && ref < references_.size(); i++, ref++) {
        float th = angles_[i] * float(M_PI) / 180.0f;
        size_t n = size_t(sample_count_);
        std::vector<float> wi(n * 3), wo(n * 3, 0.0f), u(n * 4), ev(n * 3), pdf(n), f(n * 3), wso(n * 3), eta(n),
            spdf(n);
        std::vector<int32_t> meas(n, 1), smeas(n);
        for (size_t k = 0; k < n; k++) {
          wi[3 * k] = std::sin(th);
          wi[3 * k + 1] = 0.0f;
          wi[3 * k + 2] = std::cos(th);
          u[4 * k] = rng.NextFloat();
          u[4 * k + 1] = rng.NextFloat();
          u[4 * k + 2] = extra.NextFloat();
          u[4 * k + 3] = extra.NextFloat();
        }
        CheckHk(hk_bsdf_batch(GpuContext(), records.data(), uint32_t(records.size() / HK_MATERIAL_STRIDE),
                              uint32_t(idx), nullptr, uint32_t(n), wi.data(), wo.data(), meas.data(), u.data(),
                              ev.data(), pdf.data(), f.data(), wso.data(), smeas.data(), eta.data(), spdf.data()),
                "ttest: BSDF samples");
        double mean = 0, var = 0;
        for (size_t k = 0; k < n; k++) {
          double x = double(Color3f(f[3 * k], f[3 * k + 1], f[3 * k + 2]).GetLuminance());
          double dl = x - mean;
          mean += dl / double(k + 1);
          var += dl * (x - mean);
        }
        var /= double(n - 1);
        std::string msg;
        total++;
        if (StudentsTTest(mean, var, references_[ref], long(n), significance_, int(references_.size()), msg)) passed++;
        printf("Testing (angle=%g): %s\n  %s\n", angles_[i], b->ToString().c_str(), msg.c_str());
      }
    }
  } else {
    if (references_.size() != scenes_.size())
      throw HikariException("Specified a different number of scenes and reference values!");
    for (size_t i = 0; i < scenes_.size(); i++) {
      Scene *s = scenes_[i];
      const Camera *cam = s->GetCamera();
      if (cam->width != 1 || cam->height != 1)
        throw HikariException("ttest: scene %zu has a %dx%d film; scene tests need a 1 x 1 film", i, cam->width,
                              cam->height);
      std::vector<float> out;
      if (!RenderPixelSamples(s->GetIntegrator(), s, uint32_t(i), uint32_t(sample_count_), out))
        throw HikariException("ttest: scene %zu does not use a GPU integrator", i);
      double mean = 0, var = 0;
      for (int k = 0; k < sample_count_; k++) {
        double x = double(Color3f(out[5 * size_t(k)], out[5 * size_t(k) + 1], out[5 * size_t(k) + 2]).GetLuminance());
        double dl = x - mean;
        mean += dl / double(k + 1);
        var += dl * (x - mean);
      }
      var /= double(sample_count_ - 1);
      std::string msg;
      total++;
      if (StudentsTTest(mean, var, references_[i], sample_count_, significance_, int(references_.size()), msg)) passed++;
      printf("Testing scene %zu: %s\n  %s\n", i, s->GetIntegrator()->ToString().c_str(), msg.c_str());
    }
  }
  printf("Passed %d/%d tests.\n", passed, total);
}

std::string StudentTTest::ToString() const {
  char b[128];
  snprintf(b, sizeof b, "StudentsTTest[significanceLevel = %g, sampleCount = %d]", significance_, sample_count_);
  return b;
}

REGISTER_CLASS(StudentTTest, "ttest")

}  // namespace hikari
