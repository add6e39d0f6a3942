/*
 * oracle_mlp.c -- double-precision restatement of the cost model's forward
 * pass and loss gradient (4 -> 32 ReLU -> 32 ReLU -> C, mean squared error of
 * log-times).  TEST INFRASTRUCTURE ONLY.  The model has no counterpart in the
 * reference; it is pinned by central finite differences (tests/test_oracle.py).
 *
 * Parameter layout (floats): W1[32][4] b1[32] W2[32][32] b2[32] W3[C][32] b3[C].
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define IN 4
#define HID 32

int oracle_mlp_num_params(int c) { return HID * IN + HID + HID * HID + HID + c * HID + c; }

static void fwd(const double *P, int C, const double *x, double *h1, double *h2, double *y) {
    const double *W1 = P, *b1 = W1 + HID * IN, *W2 = b1 + HID, *b2 = W2 + HID * HID, *W3 = b2 + HID, *b3 = W3 + C * HID;
    for (int k = 0; k < HID; ++k) {
        double s = b1[k];
        for (int i = 0; i < IN; ++i) s += W1[k * IN + i] * x[i];
        h1[k] = s > 0 ? s : 0;
    }
    for (int m = 0; m < HID; ++m) {
        double s = b2[m];
        for (int k = 0; k < HID; ++k) s += W2[m * HID + k] * h1[k];
        h2[m] = s > 0 ? s : 0;
    }
    for (int j = 0; j < C; ++j) {
        double s = b3[j];
        for (int m = 0; m < HID; ++m) s += W3[j * HID + m] * h2[m];
        y[j] = s;
    }
}

void oracle_mlp_forward(const double *P, int C, const double *X, int N, double *Y) {
    double h1[HID], h2[HID];
    for (int e = 0; e < N; ++e) fwd(P, C, X + (size_t)e * IN, h1, h2, Y + (size_t)e * C);
}

/* returns the loss; grad (same layout as P) receives d loss / d P */
double oracle_mlp_loss_grad(const double *P, int C, const double *X, const double *T, int N, double *grad) {
    const int np = oracle_mlp_num_params(C);
    const double *W2 = P + HID * IN + HID, *W3 = W2 + HID * HID + HID;
    double *gW1 = grad, *gb1 = gW1 + HID * IN, *gW2 = gb1 + HID, *gb2 = gW2 + HID * HID, *gW3 = gb2 + HID,
           *gb3 = gW3 + C * HID;
    memset(grad, 0, sizeof(double) * (size_t)np);
    double *y = (double *)malloc(sizeof(double) * (size_t)C), *dy = (double *)malloc(sizeof(double) * (size_t)C);
    double h1[HID], h2[HID], dh1[HID], dh2[HID], loss = 0;
    const double scale = 1.0 / ((double)N * (double)C);
    for (int e = 0; e < N; ++e) {
        const double *x = X + (size_t)e * IN, *t = T + (size_t)e * C;
        fwd(P, C, x, h1, h2, y);
        for (int j = 0; j < C; ++j) {
            const double d = y[j] - t[j];
            loss += d * d * scale;
            dy[j] = 2.0 * d * scale;
        }
        for (int j = 0; j < C; ++j) {
            gb3[j] += dy[j];
            for (int m = 0; m < HID; ++m) gW3[j * HID + m] += dy[j] * h2[m];
        }
        for (int m = 0; m < HID; ++m) {
            double s = 0;
            for (int j = 0; j < C; ++j) s += dy[j] * W3[j * HID + m];
            dh2[m] = h2[m] > 0 ? s : 0;
            gb2[m] += dh2[m];
            for (int k = 0; k < HID; ++k) gW2[m * HID + k] += dh2[m] * h1[k];
        }
        for (int k = 0; k < HID; ++k) {
            double s = 0;
            for (int m = 0; m < HID; ++m) s += dh2[m] * W2[m * HID + k];
            dh1[k] = h1[k] > 0 ? s : 0;
            gb1[k] += dh1[k];
            for (int i = 0; i < IN; ++i) gW1[k * IN + i] += dh1[k] * x[i];
        }
    }
    free(y);
    free(dy);
    return loss;
}
