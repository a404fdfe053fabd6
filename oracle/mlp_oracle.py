"""CPU (numpy, float64) reference of the router MLP — TEST INFRASTRUCTURE ONLY.

Same parameter layout as the HIP kernels: W1[64][8], b1[64], W2[64][64], b2[64],
W3[64][64], b3[64], W4[16][64], b4[16]; ReLU hidden layers; loss
L = scale * sum(mask * (out - y)^2). Used to check the GPU forward pass and
gradients; its own gradient is pinned against central finite differences.
"""
import numpy as np

IN, HID, OUT = 8, 64, 16
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def forward(flat, x):
    p = unpack(flat)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ p["W1"].T + p["b1"], 0)
    h2 = np.maximum(h1 @ p["W2"].T + p["b2"], 0)
    h3 = np.maximum(h2 @ p["W3"].T + p["b3"], 0)
    out = h3 @ p["W4"].T + p["b4"]
    return out, (x, h1, h2, h3)


def loss(flat, x, y, m, scale):
    out, _ = forward(flat, x)
    return scale * float(np.sum(m * (out - y) ** 2))


def gradient(flat, x, y, m, scale):
    p = unpack(flat)
    out, (x, h1, h2, h3) = forward(flat, x)
    d4 = 2 * scale * m * (out - y)
    d3 = (d4 @ p["W4"]) * (h3 > 0)
    d2 = (d3 @ p["W3"]) * (h2 > 0)
    d1 = (d2 @ p["W2"]) * (h1 > 0)
    g = {"W4": d4.T @ h3, "b4": d4.sum(0), "W3": d3.T @ h2, "b3": d3.sum(0),
         "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ x, "b1": d1.sum(0)}
    return np.concatenate([g[name].ravel() for name, _ in SHAPES])
