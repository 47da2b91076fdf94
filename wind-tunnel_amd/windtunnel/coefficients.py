"""Drag and lift coefficients as the reference viewer derives them from the
force of each frame (web/js/main.js:336-351).

Per frame of `steps` lattice steps with force (fx, fy, fz):
  q  = 0.5 * uin^2 * max(1, A)          A = frontal (or planform) cell count
  cd = fx / q, cl = fy / q               (main.js:338-339)
  a non-finite cd or |cd| > 1e3 means the run diverged (main.js:340)
  the first 800 steps are a settling period with no average (main.js:343-344)
  then w = min(w + steps, 2500) and the average moves by steps / w toward the
  frame's value: a running mean while w < 2500, then an exponential window of
  about 2500 steps (main.js:345-349).

The viewer keeps this state in double precision (JS numbers); so does this class.
"""
import math

SETTLE_STEPS = 800       # main.js:344
WINDOW_STEPS = 2500      # main.js:346
DIVERGED_CD = 1e3        # main.js:340


def dynamic_pressure(uin, area):
    """0.5 uin^2 max(1, A) (main.js:338)."""
    return 0.5 * float(uin) * float(uin) * max(1.0, float(area))


class ForceAverage:
    """The viewer's Cd/Cl average over frames (main.js:336-351)."""

    def __init__(self, uin, area):
        self.q = dynamic_pressure(uin, area)
        self.cd = 0.0
        self.cl = 0.0
        self.w = 0
        self.steps = 0
        self.diverged = False

    @property
    def settled(self):
        return self.steps >= SETTLE_STEPS

    def add(self, fx, fy, steps):
        """Account one frame of `steps` steps; returns (cd, cl) of the average,
        or None while settling or once the run has diverged."""
        if self.diverged or steps <= 0:
            return None
        cd, cl = float(fx) / self.q, float(fy) / self.q
        if not math.isfinite(cd) or abs(cd) > DIVERGED_CD:
            self.diverged = True
            return None
        self.steps += int(steps)
        if self.steps < SETTLE_STEPS:
            return None
        self.w = min(self.w + int(steps), WINDOW_STEPS)
        a = steps / self.w
        self.cd += (cd - self.cd) * a
        self.cl += (cl - self.cl) * a
        return self.cd, self.cl
