// Per-front update-precision classifier (mlp.hip): sizes and host entry points.
#pragma once

namespace rv {

constexpr int MLP_IN = 16, MLP_HID = 64, MLP_OUT = 3;
constexpr int MLP_PARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT;

// host arrays; x is n x MLP_IN row-major, prob n x MLP_OUT, labels in [0, MLP_OUT)
void mlp_forward(const float *par, int n, const float *x, float *prob);
void mlp_loss_grad(const float *par, int n, const float *x, const int *y, const float *class_weight,
                   float *loss, float *grad);
void mlp_train(float *par, int n, const float *x, const int *y, const float *class_weight, int epochs,
               float lr, float *final_loss);

}  // namespace rv
