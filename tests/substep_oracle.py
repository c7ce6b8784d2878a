"""The oracle behind fvmhd3d_amd.substep.Substepper (tests only)."""
from oracle.oracle import Oracle


class OracleBackend:
    def __init__(self, limiter=0, mode="f64"):
        self.o = Oracle(mode, limiter=limiter)

    def load(self, st):
        pass

    def put(self, st, names):
        pass

    def fetch(self, st, names):
        pass

    def get_active_faces(self, st):
        return len(self.o.get_active_faces(st))

    def compute_primitives(self, st):
        self.o.compute_primitives(st, False)

    def compute_timestep(self, st):
        self.o.compute_timestep(st)

    def compute_timestep_limiter(self, st):
        self.o.compute_timestep_limiter(st)

    def compute_predictor(self, st, dtmax, min_rung):
        self.o.compute_predictor(st, dtmax, min_rung)

    def compute_reconstruction(self, st):
        self.o.compute_reconstruction(st)

    def slope_limiter(self, st):
        self.o.slope_limiter(st)

    def compute_fluid_update(self, st):
        self.o.compute_fluid_update(st)
