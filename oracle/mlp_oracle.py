"""FP64 restatement of the learned projector's forward / backward pass.

TEST INFRASTRUCTURE ONLY.  The projector has no counterpart in the reference;
this file states its mathematics independently of the HIP kernels
(kgraph_amd/csrc/kg_mlp.hip):

    h = relu(W1 x + b1),  z = W2 h + b2
    L = mean_t max(0, |z_a - z_p|^2 - |z_a - z_n|^2 + margin)

where x is the standardised row (x_raw - mu) * inv_sigma; callers apply the
standardisation (kg_mlp_get_normalization) in FP32 as the GPU gather does.

`bf16=True` rounds every matrix-product operand to bfloat16 (round to
nearest even), which is what the MFMA products consume; accumulation is FP64
here (FP32 on the GPU), so the two agree to FP32 accumulation error.  With
`bf16=False` it is the exact FP64 model, which tests pin with central finite
differences.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _r(a, bf16):
    return to_bf16(a) if bf16 else np.asarray(a, dtype=np.float64)


def forward(X, W1, b1, W2, b2, bf16=True):
    H = np.maximum(_r(X, bf16) @ _r(W1, bf16).T + b1, 0.0)
    H = H.astype(np.float32) if bf16 else H  # the GPU stores H in FP32 before the next product
    Z = _r(H, bf16) @ _r(W2, bf16).T + b2
    return H, (Z.astype(np.float32) if bf16 else Z)


def loss_grad(X, W1, b1, W2, b2, margin, bf16=True):
    """X holds triplets as rows (3t, 3t+1, 3t+2). Returns (loss, dW1, db1, dW2, db2)."""
    B = X.shape[0] // 3
    H, Z = forward(X, W1, b1, W2, b2, bf16)
    Z = np.asarray(Z, dtype=np.float64)
    za, zp, zn = Z[0::3], Z[1::3], Z[2::3]
    dap = ((za - zp) ** 2).sum(1)
    dan = ((za - zn) ** 2).sum(1)
    l = dap - dan + margin
    act = (l > 0).astype(np.float64)[:, None]
    loss = (np.maximum(l, 0.0)).sum() / B
    s = 2.0 / B
    dZ = np.zeros_like(Z)
    dZ[0::3] = act * s * (zn - zp)
    dZ[1::3] = act * -s * (za - zp)
    dZ[2::3] = act * s * (za - zn)
    if bf16:
        dZ = dZ.astype(np.float32)
    dW2 = _r(dZ, bf16).T @ _r(H, bf16)
    db2 = np.asarray(dZ, np.float64).sum(0)
    dH = (_r(dZ, bf16) @ _r(W2, bf16)) * (np.asarray(H) > 0)
    if bf16:
        dH = dH.astype(np.float32)
    dW1 = _r(dH, bf16).T @ _r(X, bf16)
    db1 = np.asarray(dH, np.float64).sum(0)
    return loss, dW1, db1, dW2, db2
