"""Immersed rigid disk: the host-side body of the direct-forcing coupling.

The field update (u += chi (U_s - u) on every face inside the smoothed mask)
and the FP64 impulse reduction run on the device (`k_solid` in the HIP
library). The body itself -- mass, inertia, Newton's third law and the explicit
Euler update -- stays on the host, as it does in the reference
(`RigidDisk`, src/cpu/solid.js:52-140; `integrateSolid`,
src/gpu/solver.js:454-510).
"""
from __future__ import annotations

import math


class RigidDisk:
    """A uniform disk with prescribed density (fluid density 1).

    Same constructor fields and meaning as src/cpu/solid.js:52-92: centre
    (x, y), radius r, density ratio, velocity (vx, vy), angular velocity omega,
    `fixed` (pinned: blocks flow, never moves) and the optional added-mass
    correction.
    """

    def __init__(self, x, y, r, density=1.0, vx=0.0, vy=0.0, omega=0.0, fixed=False,
                 addedMassCorrection=False):
        self.x, self.y, self.r = float(x), float(y), float(r)
        self.density = float(density)
        self.vx, self.vy, self.omega = float(vx), float(vy), float(omega)
        self.fixed = bool(fixed)
        self.mass = self.density * math.pi * self.r * self.r
        self.inertia = 0.5 * self.mass * self.r * self.r
        self.addedMassCorrection = bool(addedMassCorrection)
        self.addedMass = math.pi * self.r * self.r
        self.effectiveMass = self.mass + (self.addedMass if addedMassCorrection else 0.0)
        self.effectiveInertia = self.inertia + (0.5 * self.addedMass * self.r * self.r
                                                if addedMassCorrection else 0.0)
        self.force = [0.0, 0.0]
        self.torque = 0.0
        self.lastImpulse = [0.0, 0.0]

    def apply_reaction(self, impulse, dt):
        """Newton's third law from the fluid impulse (fx, fy, torque), dA included
        (src/cpu/solid.js:222-229)."""
        fx, fy, tq = impulse
        self.lastImpulse = [fx, fy]
        self.force = [-fx / dt, -fy / dt]
        self.torque = -tq / dt

    def integrate(self, dt, gravity=0.0):
        """Explicit Euler under the fluid reaction plus gravity (src/cpu/solid.js:116-128)."""
        if self.fixed:
            return
        self.vx += (self.force[0] / self.effectiveMass) * dt
        self.vy += (self.force[1] / self.effectiveMass) * dt + gravity * dt
        self.omega += (self.torque / self.effectiveInertia) * dt
        self.x += self.vx * dt
        self.y += self.vy * dt

    def momentum(self):
        return [self.mass * self.vx, self.mass * self.vy]

    def kineticEnergy(self):
        return 0.5 * self.mass * (self.vx ** 2 + self.vy ** 2) + 0.5 * self.inertia * self.omega ** 2

    def as_dict(self):
        return dict(x=self.x, y=self.y, r=self.r, vx=self.vx, vy=self.vy, omega=self.omega)

    def toJSON(self):
        return dict(x=self.x, y=self.y, r=self.r, density=self.density, vx=self.vx, vy=self.vy,
                    omega=self.omega, fixed=self.fixed, mass=self.mass, force=self.force, torque=self.torque)
