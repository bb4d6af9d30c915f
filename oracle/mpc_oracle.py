"""CPU restatement of the MPC caller's wrench estimator and frame transform (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

* ``EstimatorRef`` follows examples/force_estimator.py: construction and Fibonacci directions (:6-61), the
  hypothesis batch (:82-97), the update (:99-148) and reset (:150-155). Arrays are float32 as in the reference
  and scalar factors enter as Python floats; the random rotation takes its three uniforms as an argument
  (the reference draws them with numpy.random.rand, :63-80).
* ``wrench_to_solver_frame`` follows mpc_controller.py:311-337 (pinocchio's SE3.actInv on Force, applied
  with the last joint's world placement and then with its placement relative to the joint before it), with the
  joint frames of the oracle's own kinematics (``oracle.frames``).
"""
import numpy as np

from . import oracle as O


def _act_inv_force(R, p, f, n):
    """Inverse action of the placement (R, p) on the wrench (force f, moment n about the origin)."""
    return R.T @ f, R.T @ (n - np.cross(p, f))


def wrench_to_solver_frame(plant, q, w_world):
    """World wrench [force; moment] -> the wrench the solver takes ([force; moment] in the frame of the joint
    before the last one, mpc_controller.py:311-337), float64."""
    R, o = O.frames(plant, np.asarray(q, dtype=np.float64))
    R1, p1, R0, p0 = R[-1], o[-1], R[-2], o[-2]
    Rrel, prel = R0.T @ R1, R0.T @ (p1 - p0)
    w = np.asarray(w_world, dtype=np.float64)
    f1, n1 = _act_inv_force(R1, p1, w[:3], w[3:])
    f2, n2 = _act_inv_force(Rrel, prel, f1, n1)
    return np.concatenate([f2, n2])


def fibonacci_directions(n):
    """n unit vectors on the Fibonacci spiral (force_estimator.py:33-61), float32 [n, 3]."""
    out = np.zeros((n, 3), dtype=np.float32)
    golden = (1.0 + np.sqrt(5.0)) / 2.0
    for i in range(n):
        y = 1.0 - 2.0 * i / (n - 1) if n > 1 else 0.0
        r, th = np.sqrt(1.0 - y * y), 2.0 * np.pi * i / golden
        out[i] = (r * np.cos(th), y, r * np.sin(th))
    return out


def rotation_from_uniforms(u1, u2, u3):
    """Rotation matrix (float32) of the unit quaternion drawn from three uniforms (force_estimator.py:63-80)."""
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x, y = a * np.sin(2.0 * np.pi * u2), a * np.cos(2.0 * np.pi * u2)
    z, w = b * np.sin(2.0 * np.pi * u3), b * np.cos(2.0 * np.pi * u3)
    return np.array([[1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)],
                     [2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)],
                     [2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)]], dtype=np.float32)


class EstimatorRef:
    def __init__(self, batch_size, initial_radius=10.0, min_radius=1.0, max_radius=100.0, smoothing_factor=0.3):
        assert batch_size > 3
        self.B = batch_size
        self.r_min, self.r_max, self.smoothing = min_radius, max_radius, smoothing_factor
        self.dirs = fibonacci_directions(batch_size - 3)
        self.reset(initial_radius)

    def reset(self, radius):
        self.est = np.zeros(6, dtype=np.float32)
        self.mom = np.zeros(6, dtype=np.float32)
        self.smooth = np.zeros(6, dtype=np.float32)
        self.radius = float(radius)
        self.confidence = 0.0
        self.history = []
        self.R = np.eye(3, dtype=np.float32)

    def batch(self):
        """World-frame hypotheses [B, 6] float32 (force_estimator.py:82-97)."""
        out = np.zeros((self.B, 6), dtype=np.float32)
        out[0] = self.smooth
        out[2] = self.smooth + np.float32(0.5) * self.mom
        base = np.float32(0.7) * self.smooth[:3] + np.float32(0.3) * self.est[:3]
        for i in range(3, self.B):
            d = (self.R.astype(np.float64) @ self.dirs[i - 3].astype(np.float64)).astype(np.float32)
            out[i, :3] = (base.astype(np.float64) + self.radius * d.astype(np.float64)).astype(np.float32)
            out[i, 3:] = self.smooth[3:]
        return out

    def update(self, best, errors, alpha, beta, uniforms):
        """force_estimator.py:99-148 with the rotation's uniforms given."""
        f32 = np.float32
        self.history.append(float(np.min(errors)))
        bf = self.batch()[best]
        delta = bf - self.est
        self.mom = f32(beta) * self.mom + f32(1.0 - beta) * delta
        raw = f32(alpha) * bf + f32(1.0 - alpha) * self.est
        self.est = f32(0.8) * self.est + f32(0.2) * (raw + f32(0.5) * self.mom)
        self.smooth = f32(1.0 - self.smoothing) * self.smooth + f32(self.smoothing) * self.est
        if best < 3:
            self.radius *= 0.95
            self.confidence = min(1.0, self.confidence + 0.05)
        else:
            self.radius *= 1.05
            self.confidence = max(0.0, self.confidence - 0.1)
        self.radius = float(np.clip(self.radius, self.r_min, self.r_max))
        if len(self.history) > 5:
            recent = np.array(self.history[-5:])
            if np.std(recent) < 0.01:
                self.radius *= 0.9
            elif recent[-1] > 1.5 * np.mean(recent[:-1]):
                self.radius *= 1.3
                self.confidence *= 0.5
            self.radius = float(np.clip(self.radius, self.r_min, self.r_max))
        self.R = rotation_from_uniforms(*uniforms)
