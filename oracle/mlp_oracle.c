/*
 * mlp_oracle.c -- FP64 forward/backward of the stepping network
 * (TEST INFRASTRUCTURE ONLY; the checker of voreen_amd/csrc/mlp.hip).
 *
 * Pinned by central finite differences of its own loss in FP64
 * (tests/test_learned.py::test_mlp_oracle_finite_differences) and by the
 * closed form of a network whose hidden layers are zero (output = bias).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#include "dvr_oracle.h"

#define NI 16
#define NH 64
#define NO 2

static double bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
    u = (u + r) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_d(double x) {
    return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
}

typedef struct {
    double W1[NI * NH], b1[NH], W2[NH * NH], b2[NH], W3[NH * NO], b3[NO];
} net_t;

static void unpack(const double* p, int quantize, net_t* n) {
    const double* q = p;
    for (int i = 0; i < NI * NH; ++i) n->W1[i] = quantize ? bf16_round(q[i]) : q[i];
    q += NI * NH;
    for (int i = 0; i < NH; ++i) n->b1[i] = q[i];
    q += NH;
    for (int i = 0; i < NH * NH; ++i) n->W2[i] = quantize ? bf16_round(q[i]) : q[i];
    q += NH * NH;
    for (int i = 0; i < NH; ++i) n->b2[i] = q[i];
    q += NH;
    for (int i = 0; i < NH * NO; ++i) n->W3[i] = quantize ? bf16_round(q[i]) : q[i];
    q += NH * NO;
    for (int i = 0; i < NO; ++i) n->b3[i] = q[i];
}

static void row_forward(const net_t* n, const double* x, double* z1, double* h1, double* z2,
                        double* h2, double* y) {
    for (int j = 0; j < NH; ++j) {
        double s = n->b1[j];
        for (int k = 0; k < NI; ++k) s += x[k] * n->W1[k * NH + j];
        z1[j] = s;
        h1[j] = gelu(s);
    }
    for (int j = 0; j < NH; ++j) {
        double s = n->b2[j];
        for (int k = 0; k < NH; ++k) s += h1[k] * n->W2[k * NH + j];
        z2[j] = s;
        h2[j] = gelu(s);
    }
    for (int j = 0; j < NO; ++j) {
        double s = n->b3[j];
        for (int k = 0; k < NH; ++k) s += h2[k] * n->W3[k * NO + j];
        y[j] = s;
    }
}

void dvr_mlp_forward(const float* params, int32_t quantize, const float* x, int32_t rows, double* y) {
    double* p = (double*)malloc(sizeof(double) * DVR_MLP_PARAMS);
    for (int i = 0; i < DVR_MLP_PARAMS; ++i) p[i] = params[i];
    net_t* n = (net_t*)malloc(sizeof(net_t));
    unpack(p, quantize, n);
    for (int r = 0; r < rows; ++r) {
        double xr[NI], z1[NH], h1[NH], z2[NH], h2[NH];
        for (int k = 0; k < NI; ++k) xr[k] = bf16_round(x[(size_t)r * NI + k]);
        row_forward(n, xr, z1, h1, z2, h2, y + (size_t)r * NO);
    }
    free(n);
    free(p);
}

double dvr_mlp_loss_grad(const double* params, int32_t quantize, const float* x, const float* t,
                         int32_t rows, double* grad) {
    net_t* n = (net_t*)malloc(sizeof(net_t));
    unpack(params, quantize, n);
    memset(grad, 0, sizeof(double) * DVR_MLP_PARAMS);
    double* gW1 = grad;
    double* gb1 = gW1 + NI * NH;
    double* gW2 = gb1 + NH;
    double* gb2 = gW2 + NH * NH;
    double* gW3 = gb2 + NH;
    double* gb3 = gW3 + NH * NO;
    double loss = 0.0;
    const double inv = 1.0 / rows;
    for (int r = 0; r < rows; ++r) {
        double xr[NI], z1[NH], h1[NH], z2[NH], h2[NH], y[NO];
        for (int k = 0; k < NI; ++k) xr[k] = bf16_round(x[(size_t)r * NI + k]);
        row_forward(n, xr, z1, h1, z2, h2, y);
        const float* tr = t + (size_t)r * 3;
        const double s = 1.0 / (1.0 + exp(-y[0]));
        double dy[NO];
        dy[0] = (s - tr[0]) * inv;
        dy[1] = 2.0 * tr[2] * (y[1] - tr[1]) * inv;
        loss += (fmax(y[0], 0.0) - y[0] * tr[0] + log1p(exp(-fabs(y[0])))) * inv;
        loss += tr[2] * (y[1] - tr[1]) * (y[1] - tr[1]) * inv;
        double dz2[NH], dz1[NH];
        for (int j = 0; j < NO; ++j) gb3[j] += dy[j];
        for (int k = 0; k < NH; ++k) {
            double dh = 0.0;
            for (int j = 0; j < NO; ++j) {
                gW3[k * NO + j] += h2[k] * dy[j];
                dh += dy[j] * n->W3[k * NO + j];
            }
            dz2[k] = dh * gelu_d(z2[k]);
            gb2[k] += dz2[k];
        }
        for (int k = 0; k < NH; ++k) {
            double dh = 0.0;
            for (int j = 0; j < NH; ++j) {
                gW2[k * NH + j] += h1[k] * dz2[j];
                dh += dz2[j] * n->W2[k * NH + j];
            }
            dz1[k] = dh * gelu_d(z1[k]);
            gb1[k] += dz1[k];
        }
        for (int f = 0; f < NI; ++f)
            for (int j = 0; j < NH; ++j) gW1[f * NH + j] += xr[f] * dz1[j];
    }
    free(n);
    return loss;
}
