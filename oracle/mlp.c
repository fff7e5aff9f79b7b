/* Decode-cost model checker.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no cost
 * model; this file is the double-precision statement of the product's
 * 5-32-32-1 ReLU MLP (parameter layout W1[32][5] | b1[32] | W2[32][32] |
 * b2[32] | W3[32] | b3) and of the gradient of its mean squared loss
 * L = (1/n) sum 0.5 (f(x_i) - y_i)^2, against which the device's training
 * kernels are checked.  The gradient itself is pinned by finite differences
 * of oracle_mlp_loss in the tests.
 */
#include <string.h>

#include "oracle.h"

#define IN 5
#define HID 32
#define OW1 0
#define OB1 (OW1 + HID * IN)
#define OW2 (OB1 + HID)
#define OB2 (OW2 + HID * HID)
#define OW3 (OB2 + HID)
#define OB3 (OW3 + HID)

static double forward(const double *p, const double *x, double *h1, double *h2)
{
    for (int i = 0; i < HID; ++i) {
        double a = p[OB1 + i];
        for (int j = 0; j < IN; ++j)
            a += p[OW1 + i * IN + j] * x[j];
        h1[i] = a > 0 ? a : 0;
    }
    for (int i = 0; i < HID; ++i) {
        double a = p[OB2 + i];
        for (int j = 0; j < HID; ++j)
            a += p[OW2 + i * HID + j] * h1[j];
        h2[i] = a > 0 ? a : 0;
    }
    double y = p[OB3];
    for (int j = 0; j < HID; ++j)
        y += p[OW3 + j] * h2[j];
    return y;
}

double oracle_mlp_predict(const double *params, const float *x)
{
    double xd[IN], h1[HID], h2[HID];
    for (int j = 0; j < IN; ++j)
        xd[j] = x[j];
    return forward(params, xd, h1, h2);
}

double oracle_mlp_loss(const double *params, const float *X, const float *Y, size_t n)
{
    double loss = 0;
    for (size_t s = 0; s < n; ++s) {
        const double e = oracle_mlp_predict(params, X + s * IN) - Y[s];
        loss += 0.5 * e * e;
    }
    return n ? loss / (double)n : 0;
}

double oracle_mlp_loss_grad(const double *p, const float *X, const float *Y, size_t n,
                            double *grad)
{
    memset(grad, 0, sizeof(double) * ORACLE_MLP_PARAMS);
    double loss = 0;
    for (size_t s = 0; s < n; ++s) {
        double x[IN], h1[HID], h2[HID], d1[HID], d2[HID];
        for (int j = 0; j < IN; ++j)
            x[j] = X[s * IN + j];
        const double e = forward(p, x, h1, h2) - Y[s];
        loss += 0.5 * e * e;
        for (int i = 0; i < HID; ++i)
            d2[i] = h2[i] > 0 ? e * p[OW3 + i] : 0;
        for (int i = 0; i < HID; ++i) {
            double a = 0;
            for (int k = 0; k < HID; ++k)
                a += p[OW2 + k * HID + i] * d2[k];
            d1[i] = h1[i] > 0 ? a : 0;
        }
        for (int i = 0; i < HID; ++i) {
            for (int j = 0; j < IN; ++j)
                grad[OW1 + i * IN + j] += d1[i] * x[j];
            grad[OB1 + i] += d1[i];
            for (int j = 0; j < HID; ++j)
                grad[OW2 + i * HID + j] += d2[i] * h1[j];
            grad[OB2 + i] += d2[i];
            grad[OW3 + i] += e * h2[i];
        }
        grad[OB3] += e;
    }
    if (n) {
        for (int k = 0; k < ORACLE_MLP_PARAMS; ++k)
            grad[k] /= (double)n;
        loss /= (double)n;
    }
    return loss;
}
