"""SPH kernel functions with the reference's signatures (src/kernels.py:21-52),
evaluated on torch tensors.  The hot path evaluates them inside the HIP
gathers; these are for callers that work on materialised pair lists."""
import numpy as np
import torch


def wendland(q, h):
    C = 7 / np.pi
    return torch.pow(1.0 - q, 4) * (1.0 + 4.0 * q) * C / h ** 2


def wendlandGrad(q, r, h):
    C = 7 / np.pi
    return -r * C / h ** 3 * (20.0 * q * (1.0 - q) ** 3)[:, None]


def spikyGrad(q, r, h):
    return -r * 30 / np.pi / h ** 3 * ((1 - q) ** 2)[:, None]


def kernel(q, h):
    return wendland(q, h)


def kernelGradient(q, r, h):
    return wendlandGrad(q, r, h)


def getKernelFunctions(kernel_name):
    if kernel_name == "wendland2":
        return wendland, wendlandGrad
    raise ValueError(f"unknown kernel {kernel_name}")
