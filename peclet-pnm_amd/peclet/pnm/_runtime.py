"""Pin the process's HIP runtime before libpnm_hip.so is loaded.

torch ships its own libamdhip64 / libhsa-runtime64. If libpnm_hip.so were loaded first it would
pull the system ROCm runtime, and torch would later load a second HIP runtime into the same
process, whose initialisation then fails to find the GPU. Importing torch first (when it is
installed) makes libpnm_hip.so bind to the runtime torch already loaded (same SONAME), so a
process always holds exactly one HIP runtime. Without torch, the system runtime is used.
"""


def pin_hip_runtime() -> None:
    try:
        import torch  # noqa: F401
    except Exception:  # torch absent or broken: the system ROCm runtime is the only one
        pass
