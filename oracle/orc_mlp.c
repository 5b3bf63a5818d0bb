/*
 * FP64 oracle: the radial-integral surrogate network (1 -> 64 -> 64 -> n_out,
 * SiLU) used in place of the GTO radial integral. TEST INFRASTRUCTURE ONLY
 * (see oracle.h).
 *
 * The reference has no learned radial integral: this network fills the slot of
 * its SoapRadialIntegral trait (featomic/src/calculators/soap/radial_integral/
 * mod.rs:19-58) and is trained on the exact GTO integral of
 * soap/radial_integral/gto.rs:115-205 (orc_radial_compute). The functions below
 * restate the network definition, its mean-squared-error loss and the
 * loss gradient with respect to every parameter, in FP64, as the checker of
 * the device training and inference kernels. They are pinned by central finite
 * differences and by an independent float64 automatic differentiation in
 * tests/test_oracle_mlp.py.
 *
 * Parameter layout (shared with the device library):
 *   W1[64], b1[64], W2[64][64] (row = output unit), b2[64],
 *   W3[n_out][64], b3[n_out]
 * Network: x -> a1 = silu(W1 x + b1) -> a2 = silu(W2 a1 + b2) -> y = W3 a2 + b3,
 * silu(z) = z / (1 + exp(-z)).
 * Loss: mean over samples and outputs of (y - t)^2.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define H ORC_MLP_HIDDEN

int64_t orc_mlp_n_params(int n_out) { return (int64_t)H + H + H * H + H + (int64_t)n_out * H + n_out; }

static double sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

/* one sample forward; keeps the pre-activations for the backward pass */
static void forward_one(int n_out, const double* p, double x, double* z1, double* a1, double* z2, double* a2,
                        double* y) {
    const double* W1 = p;
    const double* b1 = W1 + H;
    const double* W2 = b1 + H;
    const double* b2 = W2 + H * H;
    const double* W3 = b2 + H;
    const double* b3 = W3 + (size_t)n_out * H;
    for (int u = 0; u < H; u++) {
        z1[u] = W1[u] * x + b1[u];
        a1[u] = z1[u] * sigmoid(z1[u]);
    }
    for (int u = 0; u < H; u++) {
        double s = b2[u];
        for (int k = 0; k < H; k++) s += W2[u * H + k] * a1[k];
        z2[u] = s;
        a2[u] = s * sigmoid(s);
    }
    for (int o = 0; o < n_out; o++) {
        double s = b3[o];
        for (int k = 0; k < H; k++) s += W3[(size_t)o * H + k] * a2[k];
        y[o] = s;
    }
}

void orc_mlp_forward(int n_out, int64_t n_samples, const double* params, const double* inputs, double* outputs) {
    double z1[H], a1[H], z2[H], a2[H];
    for (int64_t s = 0; s < n_samples; s++) {
        forward_one(n_out, params, inputs[s], z1, a1, z2, a2, outputs + s * n_out);
    }
}

/* d silu / dz = sigma(z) (1 + z (1 - sigma(z))) */
static double dsilu(double z) {
    const double sg = sigmoid(z);
    return sg * (1.0 + z * (1.0 - sg));
}

double orc_mlp_loss_gradient(int n_out, int64_t n_samples, const double* params, const double* inputs,
                             const double* targets, double* gradient) {
    const int64_t np = orc_mlp_n_params(n_out);
    memset(gradient, 0, (size_t)np * sizeof(double));
    double* gW1 = gradient;
    double* gb1 = gW1 + H;
    double* gW2 = gb1 + H;
    double* gb2 = gW2 + H * H;
    double* gW3 = gb2 + H;
    double* gb3 = gW3 + (size_t)n_out * H;
    const double* W2 = params + 2 * H;
    const double* W3 = W2 + H * H + H;
    double* y = malloc((size_t)n_out * sizeof(double));
    double* dy = malloc((size_t)n_out * sizeof(double));
    const double scale = 2.0 / ((double)n_samples * n_out);
    double loss = 0.0;
    for (int64_t s = 0; s < n_samples; s++) {
        double z1[H], a1[H], z2[H], a2[H], da2[H], dz2[H], da1[H];
        const double x = inputs[s];
        forward_one(n_out, params, x, z1, a1, z2, a2, y);
        for (int o = 0; o < n_out; o++) {
            const double d = y[o] - targets[s * n_out + o];
            loss += d * d;
            dy[o] = scale * d;
        }
        for (int k = 0; k < H; k++) da2[k] = 0.0;
        for (int o = 0; o < n_out; o++) {
            gb3[o] += dy[o];
            for (int k = 0; k < H; k++) {
                gW3[(size_t)o * H + k] += dy[o] * a2[k];
                da2[k] += W3[(size_t)o * H + k] * dy[o];
            }
        }
        for (int u = 0; u < H; u++) dz2[u] = da2[u] * dsilu(z2[u]);
        for (int k = 0; k < H; k++) da1[k] = 0.0;
        for (int u = 0; u < H; u++) {
            gb2[u] += dz2[u];
            for (int k = 0; k < H; k++) {
                gW2[u * H + k] += dz2[u] * a1[k];
                da1[k] += W2[u * H + k] * dz2[u];
            }
        }
        for (int u = 0; u < H; u++) {
            const double dz1 = da1[u] * dsilu(z1[u]);
            gW1[u] += dz1 * x;
            gb1[u] += dz1;
        }
    }
    free(y);
    free(dy);
    return loss / ((double)n_samples * n_out);
}
