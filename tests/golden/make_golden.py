"""Regenerate the golden vectors in this folder from the reference Velox CPU
library (strict floating point), built from /root/reference by oracle/ref.mk.

    python tests/golden/make_golden.py

Writes one folder per case: input records, flat geometry, and reference
outputs (see oracle/ref_harness.cpp for what each case pins).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("the reference sources are not present; fixtures cannot be regenerated here")
    subprocess.check_call(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "harness", "-j8"], cwd=ROOT)
    for name in os.listdir(HERE):
        path = os.path.join(HERE, name)
        if os.path.isdir(path):
            shutil.rmtree(path)
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixtures", HERE], cwd=ROOT)
    # The full 300-step 1,000-box lattice (several minutes of strict CPU).
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "lattice1000"],
                          cwd=ROOT)
    # Sleeping on: towers fall asleep island by island, a late strike wakes one.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "settle"], cwd=ROOT)
    # 76 joints (chains of ball joints and one of most other types), with breaking;
    # joints_trig: the atan2f/acosf joint types beside 64 chain joints.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "joints"], cwd=ROOT)
    # MotionQuality::High bullets and boxes through the multi-TOI scheduler.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "hightoi"], cwd=ROOT)
    # World::queryMultiToi for every dynamic body of the same scene at steps 5, 12 and 20.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "toiquery"], cwd=ROOT)
    # Origin shifts in the middle of a lattice and of a heightfield terrain trajectory.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "shift"], cwd=ROOT)
    # The crowded regime of a large pile from step 1: a 35,000-contact island,
    # ~700-900 colours (states, counts and the last step's events only).
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "slabstack"],
                          cwd=ROOT)
    # The CCD stress scene reduced to a size the strict CPU steps in seconds:
    # 120 steps with the deepest contact and total energy of every step.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "ccdsmall"], cwd=ROOT)
    # Sphere pairs: the CPU backend's analytic sphere-sphere contacts next to
    # the generic collidePair contacts of the same pairs.
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "spheres"], cwd=ROOT)
    # The 1,000-box lattice with sleeping on over 300 steps (several minutes).
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures"), "fixture", HERE, "lattice1000_sleep"],
                          cwd=ROOT)


if __name__ == "__main__":
    main()
