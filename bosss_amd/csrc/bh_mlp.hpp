// Learned per-cell correction network (see bh_mlp.hip).
#pragma once
#include "bh_internal.hpp"

namespace bh {

// Parameters are one flat FP32 array (the master copy), in this order:
//   W1 [hid][in_p], b1 [hid], W2 [hid][hid], b2 [hid], W3 [out_p][hid], b3 [out_p]
// (W rows = output units; padded input columns / output rows are zero).
struct Mlp {
    bh_ctx* ctx;
    int in, hid, out, in_p, out_p;
    int nw = 0;
    int step = 0;
    DevBuf params, grad, m1, m2, wbf, part, lossp;
    Mlp(bh_ctx* c, int in, int hid, int out, uint64_t seed);
    size_t off_W1() const { return 0; }
    size_t off_b1() const { return (size_t)hid * in_p; }
    size_t off_W2() const { return off_b1() + hid; }
    size_t off_b2() const { return off_W2() + (size_t)hid * hid; }
    size_t off_W3() const { return off_b2() + hid; }
    size_t off_b3() const { return off_W3() + (size_t)out_p * hid; }
    void set_params(const float* host);
    void refresh_bf16();
    void infer(int J, const float* X, float* Y);                          // Y[J][out_p]
    static int gradient_blocks(int B) { return B < 1 ? 1 : (B < 256 ? B : 256); }
    void launch_gradient(int B, int nb, const int* idx, const float* X, const float* Tg, float* part, float* lossp,
                         bool upstream);
    double gradient(int B, const int* idx, const float* X, const float* Tg);  // -> grad (device), returns loss
    // the same gradient with its partials in the fixed per-GPU workspace; no
    // allocation and no host synchronisation (grad_out, loss_dev: device, may be null)
    void gradient_noalloc(int B, const float* X, const float* Tg, float* grad_out, double* loss_dev);
    // parameter gradient (-> grad) of a loss defined outside the network, from dL/dY [B][out_p]
    void backward(int B, const float* X, const float* dY);
    void adam(float lr);
};

void mlp_noalloc_setup(int device);   // per device, from bh_init
size_t mlp_noalloc_capacity();        // floats in the allocation-free workspace
void mlp_features(bh_op* op, const void* r, int prec, float* X, float* s, int in_p);
void mlp_correction(bh_op* op, const float* s, const float* Y, int out_p, void* e, int prec);
void mlp_targets(bh_op* op, const float* s, const double* e, int out_p, float* Tg);
void mlp_upstream(bh_op* op, const float* s, const double* g, double c, int out_p, float* dY);

}  // namespace bh

struct bh_mlp {
    bh::Mlp* m = nullptr;
};
