"""Write tests/golden/slcg_profile.txt: the output of the reference's own
inc/slcgProfile.cpp, compiled unchanged into oracle/_ref/slcg_profile by
`make -C oracle ref`.

The program seeds a 64-bit SLCG with time(0), prints the low bit of the first
20 states, then steps the base generator (SLCGen, inc/kmcRandom.h:122-124)
2^30 times and prints, for the folded word t = lo ^ hi of every state, the
count of each of its 32 bits and two 12-bucket histograms: of the top byte
(t >> 24) and of the 23-bit randomFloat (t & KMC_M_LCG_RAND_RED) / 2^23
(inc/slcgProfile.cpp:30-60, inc/kmcRandom.h:54-55).  Its seed is printed, so
the fixture is self-contained; tests/test_oracle_slcg_profile.py replays the
same 2^30 draws through the oracle's fold and mask and must reproduce every
number.  Runs in the development container only (about 30 s):

    python tests/golden/make_slcg_profile.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "..", "oracle")


def main():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "ref"])
    out = subprocess.check_output([os.path.join(ORACLE, "_ref", "slcg_profile")], text=True)
    with open(os.path.join(HERE, "slcg_profile.txt"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
