"""MemoryReport / ThroughputReport, src/memory_report.mojo:36-191."""
from __future__ import annotations

from dataclasses import dataclass


def _fmt_bytes(b: float) -> str:
    for unit in ("B", "KB", "MB", "GB", "TB"):
        if abs(b) < 1024.0 or unit == "TB":
            return f"{b:.1f} {unit}" if unit != "B" else f"{int(b)} B"
        b /= 1024.0
    return f"{b:.1f} TB"


@dataclass
class MemoryReport:
    rk_stage_bytes: int
    dg_operators_bytes: int
    limiter_bytes: int
    mesh_connectivity_bytes: int
    halo_device_bytes: int = 0
    halo_pinned_bytes: int = 0

    def total_device_bytes(self) -> int:
        return (self.rk_stage_bytes + self.dg_operators_bytes + self.limiter_bytes + self.mesh_connectivity_bytes
                + self.halo_device_bytes)

    def print(self) -> None:
        rows = [("RK stage buffers", self.rk_stage_bytes), ("DG operators", self.dg_operators_bytes),
                ("Limiter scratch", self.limiter_bytes), ("Mesh connectivity", self.mesh_connectivity_bytes),
                ("Halo (device)", self.halo_device_bytes), ("Halo (pinned host)", self.halo_pinned_bytes)]
        print("Device memory:")
        for name, v in rows:
            print(f"  {name:<22s} {_fmt_bytes(v):>12s}")
        print(f"  {'Total device':<22s} {_fmt_bytes(self.total_device_bytes()):>12s}")


@dataclass
class ThroughputReport:
    """The reference's four fields, plus device-side figures from HIP events
    around the measured steps (0 when not measured): the time per stage,
    the state bytes per stage over it against the 8 TB/s HBM peak, and for
    the FP32-MFMA form the dense contraction flops per stage over it
    against the 157.3 TFLOP/s FP32 matrix peak."""
    steps: int
    wall_seconds: float
    dof: int
    state_bytes_per_step: int
    stage_seconds: float = 0.0
    achieved_hbm_gbs: float = 0.0
    hbm_fraction: float = 0.0
    mfma_tflops: float = 0.0
    mfma_utilization: float = 0.0

    def dof_per_second(self) -> float:
        return self.dof * self.steps / self.wall_seconds

    def dof_stage_updates_per_second(self) -> float:
        return 3.0 * self.dof_per_second()

    def state_bandwidth_gbs(self) -> float:
        return self.state_bytes_per_step * self.steps / self.wall_seconds * 1e-9

    def print(self) -> None:
        print(f"Throughput: {self.steps} steps in {self.wall_seconds * 1e3:.3f} ms "
              f"({self.wall_seconds / self.steps * 1e3:.4f} ms/step)")
        print(f"  {self.dof_per_second() / 1e9:.3f} GDOF/s, {self.dof_stage_updates_per_second():.4e} "
              f"DOF-stage-updates/s, state >= {self.state_bandwidth_gbs():.1f} GB/s")
        if self.stage_seconds > 0.0:
            print(f"  device: {self.stage_seconds * 1e6:.2f} us/stage, {self.achieved_hbm_gbs:.0f} GB/s "
                  f"({self.hbm_fraction * 100:.1f} % of HBM)"
                  + (f", MFMA {self.mfma_tflops:.2f} TFLOP/s ({self.mfma_utilization * 100:.1f} % of the "
                     f"FP32 matrix peak)" if self.mfma_tflops > 0.0 else ""))
