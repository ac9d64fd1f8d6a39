"""MI355X-native per-iteration path of a structured multi-block finite-volume
turbomachinery solver: host-side mirror of the reference's CSpaceDiscr /
CTimeDiscr / CMesh / CSolver surface over the libfv3d C ABI."""
