"""Extract the integer shift tables of the reference's DM plan file as a
test fixture.

Reads ``configs/dm_plan.npz`` of dsa110-rt as data (``allow_pickle=False``;
written by ``tools/build_dm_plan.py:322-387``) and stores its DM trial tables,
the three shift tables and the two sample periods from its metadata in
``tests/golden/dm_plan_tables.npz``.  Run once where the reference is
checked out:

    python tests/golden/make_dm_plan_fixture.py /root/reference
"""
import json
import sys
from pathlib import Path

import numpy as np

KEYS = ("fine_dm", "coarse_dm", "fine_to_coarse", "time_shift_corr_stage1",
        "time_shift_corr_stage2", "time_shift_search")


def main(ref_root: str) -> None:
    src = Path(ref_root) / "configs" / "dm_plan.npz"
    z = np.load(src, allow_pickle=False)
    meta = json.loads(str(z["metadata"]))
    out = Path(__file__).with_name("dm_plan_tables.npz")
    np.savez_compressed(out, **{k: z[k] for k in KEYS},
                        t_int_fast_us=np.float64(meta["t_int_fast_us"]),
                        t_int_search_us=np.float64(meta["t_int_search_us"]))
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
