"""CPU restatement of the reference's PLY export (src/io/ply.js:12-100) and
of Session.exportPlyBlob's camera centres (src/session.js:32-38, :637-645),
in the reference's arithmetic: JS numbers (f64) over Float32Array inputs,
f32 stores. Test infrastructure: only tests/ use it, for the expected results
of the trainer's gs_export_ply.

Math.exp / Math.log / Math.sqrt come from Python's math module; Math.hypot
is taken as the square root of the sum of squares (V8 scales and
compensates that sum; the two agree to an f64 ulp, below the f32 stores
except at rounding ties)."""
import math
import struct

import numpy as np

STRIDE = 16
SH_C0 = 0.28209479177387814


def cam_position(R, t):
    """session.js:32-38: -(R^T t)."""
    return [-(R[a] * t[0] + R[3 + a] * t[1] + R[6 + a] * t[2]) for a in range(3)]


def bake_opacity_compensation(data, n, f, cam_pos):
    """ply.js:16-37; data: n x 16 float32, cam_pos: flat float32 array."""
    out = np.array(data, np.float32).reshape(-1, STRIDE).copy()
    cp = np.asarray(cam_pos, np.float32)
    nc = len(cp) // 3
    cpl = [float(v) for v in cp]
    for i in range(n):
        b = [float(v) for v in data[i, :14]]
        z2min = math.inf
        for c in range(nc):
            dx = b[0] - cpl[3 * c]
            dy = b[1] - cpl[3 * c + 1]
            dz = b[2] - cpl[3 * c + 2]
            d2 = dx * dx + dy * dy + dz * dz
            if d2 < z2min:
                z2min = d2
        z = max(1e-3, math.sqrt(z2min))
        s_mean = math.exp((b[3] + b[4] + b[5]) / 3)
        s2d = f * s_mean / z
        comp = (s2d * s2d) / (s2d * s2d + 0.3)
        opa = comp / (1 + math.exp(-b[13]))
        cl = min(1 - 1e-6, max(1e-6, opa))
        out[i, 13] = np.float32(math.log(cl / (1 - cl)))
    return out


def gaussians_to_ply(data, n, sh=None, shK=0):
    """ply.js:47-100 -> bytes."""
    K = shK if sh is not None else 0
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    props += [f"f_rest_{k}" for k in range(3 * K)]
    props += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    header = ("ply\nformat binary_little_endian 1.0\n" + f"element vertex {n}\n"
              + "\n".join(f"property float {p}" for p in props) + "\nend_header\n").encode()
    rows = np.zeros((n, len(props)), np.float32)

    def sig(v):
        return 1 / (1 + math.exp(-v))

    for i in range(n):
        d = [float(v) for v in data[i]]
        rows[i, 0:3] = data[i, 0:3]
        for c in range(3):
            rows[i, 6 + c] = np.float32((sig(d[10 + c]) - 0.5) / SH_C0)
        if K:
            rows[i, 9:9 + 3 * K] = sh[i]
        q = 9 + 3 * K
        rows[i, q] = data[i, 13]
        rows[i, q + 1:q + 4] = data[i, 3:6]
        qw, qx, qy, qz = d[6], d[7], d[8], d[9]
        ql = math.sqrt(qw * qw + qx * qx + qy * qy + qz * qz)
        if ql < 1e-6:
            qw, qx, qy, qz = 1.0, 0.0, 0.0, 0.0
        else:
            qw, qx, qy, qz = qw / ql, qx / ql, qy / ql, qz / ql
        rows[i, q + 4:q + 8] = np.array([qw, qx, qy, qz], np.float64).astype(np.float32)
    return header + rows.astype("<f4").tobytes()
