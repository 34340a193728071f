// C ABI of the learned row binning (include/sdl_device.h): predictor
// lifetime, parameters, training, gradients, and predicted symbolic plans.
#include "btda_plan.h"
#include "mlp.h"

#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

struct sdl_mlp
{
    sdl::RowPredictor R;
    sdl::MlpScratch S;
};

namespace {
const char MAGIC[8] = {'S', 'D', 'L', 'M', 'L', 'P', '1', '\0'};
} // namespace

extern "C" {
#pragma GCC visibility push(default)

int sdl_mlp_create(uint64_t seed, sdl_mlp **out)
{
    if (!out)
    {
        sdl::set_error("sdl_mlp_create: null output");
        return 1;
    }
    sdl_mlp *p = new (std::nothrow) sdl_mlp();
    if (!p)
    {
        sdl::set_error("sdl_mlp_create: out of host memory");
        return 1;
    }
    sdl::mlp_init(p->R, seed);
    *out = p;
    return 0;
}

void sdl_mlp_destroy(sdl_mlp *p)
{
    if (!p) return;
    if (sdl::g_row_predictor == &p->R) sdl::g_row_predictor = nullptr;
    sdl::sync();
    delete p;
}

int sdl_mlp_num_params(void) { return sdl::MLP_NPARAM; }

int sdl_mlp_get_params(const sdl_mlp *p, float *h)
{
    if (!p || !h) return 1;
    sdl::sync();
    sdl::d2h(h, p->R.params.p, sdl::MLP_NPARAM * sizeof(float));
    return 0;
}

int sdl_mlp_set_params(sdl_mlp *p, const float *h)
{
    if (!p || !h) return 1;
    sdl::h2d(p->R.params.p, h, sdl::MLP_NPARAM * sizeof(float));
    sdl::sync();
    return 0;
}

int sdl_mlp_set_tau(sdl_mlp *p, float tau)
{
    if (!p || !(tau > 0.0f && tau <= 1.0f))
    {
        sdl::set_error("sdl_mlp_set_tau: tau must be in (0, 1]");
        return 1;
    }
    p->R.tau = tau;
    return 0;
}

int sdl_mlp_forward(const sdl_mlp *p, const float *dX, long long n, float *dZ)
{
    if (!p || n < 0) return 1;
    sdl::mlp_forward(p->R, dX, n, dZ, sdl::stream());
    sdl::sync();
    return 0;
}

int sdl_mlp_predict(const sdl_mlp *p, const float *dX, long long n, unsigned char *dbins)
{
    if (!p || n < 0) return 1;
    sdl::mlp_predict(p->R, dX, n, dbins, sdl::stream());
    sdl::sync();
    return 0;
}

int sdl_mlp_gradients(sdl_mlp *p, const float *dX, const int *dY, long long n, float *dG,
                      float *loss)
{
    if (!p || n <= 0 || !dG)
    {
        sdl::set_error("sdl_mlp_gradients: invalid arguments");
        return 1;
    }
    float l = sdl::mlp_gradients(p->R, dX, dY, n, dG, p->S, sdl::stream());
    if (loss) *loss = l;
    return 0;
}

int sdl_mlp_train(sdl_mlp *p, const float *dX, const int *dY, long long n, int epochs, int batch,
                  float lr, uint64_t seed, float *final_loss)
{
    if (!p || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f))
    {
        sdl::set_error("sdl_mlp_train: invalid arguments");
        return 1;
    }
    float l = sdl::mlp_train(p->R, p->S, dX, dY, n, epochs, batch, lr, seed, sdl::stream());
    if (final_loss) *final_loss = l;
    return 0;
}

int sdl_mlp_save(const sdl_mlp *p, const char *path)
{
    if (!p || !path) return 1;
    std::vector<float> h(sdl::MLP_NPARAM);
    sdl::sync();
    sdl::d2h(h.data(), p->R.params.p, h.size() * sizeof(float));
    uint64_t sum = sdl::fnv1a64(h.data(), h.size() * sizeof(float));
    FILE *f = std::fopen(path, "wb");
    if (!f)
    {
        sdl::set_error("sdl_mlp_save: cannot open %s", path);
        return 1;
    }
    int32_t np = sdl::MLP_NPARAM;
    bool ok = std::fwrite(MAGIC, 1, 8, f) == 8 && std::fwrite(&np, 4, 1, f) == 1 &&
              std::fwrite(&p->R.tau, 4, 1, f) == 1 &&
              std::fwrite(h.data(), 4, h.size(), f) == h.size() && std::fwrite(&sum, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) sdl::set_error("sdl_mlp_save: write failed for %s", path);
    return ok ? 0 : 1;
}

int sdl_mlp_load(sdl_mlp *p, const char *path)
{
    if (!p || !path) return 1;
    FILE *f = std::fopen(path, "rb");
    if (!f)
    {
        sdl::set_error("sdl_mlp_load: cannot open %s", path);
        return 1;
    }
    char magic[8];
    int32_t np = 0;
    float tau = 0.0f;
    std::vector<float> h(sdl::MLP_NPARAM);
    uint64_t sum = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, MAGIC, 8) == 0 &&
              std::fread(&np, 4, 1, f) == 1 && np == sdl::MLP_NPARAM &&
              std::fread(&tau, 4, 1, f) == 1 && std::fread(h.data(), 4, h.size(), f) == h.size() &&
              std::fread(&sum, 8, 1, f) == 1;
    std::fclose(f);
    if (!ok || sum != sdl::fnv1a64(h.data(), h.size() * sizeof(float)) || !(tau > 0.0f && tau <= 1.0f))
    {
        sdl::set_error("sdl_mlp_load: %s is not a valid predictor file (format or checksum)", path);
        return 1;
    }
    p->R.tau = tau;
    sdl::h2d(p->R.params.p, h.data(), h.size() * sizeof(float));
    sdl::sync();
    return 0;
}

int sdl_set_row_predictor(const sdl_mlp *p)
{
    sdl::g_row_predictor = p ? &p->R : nullptr;
    return 0;
}

int sdl_btda_plan_create_predicted(int m, int nA, int nB, const int *dAp, const int *dAi,
                                   const int *dBp, const int *dBi, const sdl_mlp *mlp,
                                   sdl_btda_plan **out)
{
    return sdl_btda_plan_create_mode(m, nA, nB, dAp, dAi, dBp, dBi, sdl::SymMode::Predicted,
                                     mlp ? &mlp->R : nullptr, out);
}

int sdl_btda_plan_features(const sdl_btda_plan *plan, float *dF)
{
    if (!plan || !dF) return 1;
    sdl::product_row_features(plan->P.L, plan->P.A, dF, sdl::stream());
    sdl::sync();
    return 0;
}

int sdl_btda_plan_row_bins(const sdl_btda_plan *plan, int *dY)
{
    if (!plan || !dY) return 1;
    sdl::mlp_bins_of_counts(plan->P.Cp.p, plan->P.m_out, dY, sdl::stream());
    sdl::sync();
    return 0;
}

int sdl_btda_plan_pred_stats(const sdl_btda_plan *plan, long long *o)
{
    if (!plan || !o) return 1;
    const sdl::SymStats &s = plan->P.stats;
    o[0] = s.predicted_rows;
    o[1] = s.overflow_rows;
    o[2] = s.misbin_rows;
    o[3] = s.exact_rows;
    o[4] = s.candidate_rows;
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
