"""proxpde_amd: MI355X-native element assembly and operator apply for the
proxpde Builder path (C ABI in include/proxpde_amd.h, bindings in .native)."""

__all__ = ["native", "workloads"]
