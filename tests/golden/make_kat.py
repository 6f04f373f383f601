"""Writes tests/golden/rng_kat.json: published known-answer vectors for the
two integer generators the reference seeds its particles with
(src/Collision.h:11-16 splitmix64, src/Collision.mm:21-27 PCG-XSH-RR).

The numbers are copied from their publications, not computed here:
  * splitmix64, seed 1234567, first five outputs -- Rosetta Code task
    "Pseudo-random numbers/Splitmix64" (Vigna's reference generator);
  * pcg32, pcg32_srandom_r(&rng, 42u, 54u), first six pcg32_random_r outputs
    -- M. O'Neill, pcg-c-basic pcg32-demo "Round 1" output.
  * splitmix64, seed 0, first output 0xe220a8397b1dcdaf (the same generator's
    widely quoted first value).
"""
import json
import os

KAT = {
    "splitmix64_seed_1234567": [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                4593380528125082431, 16408922859458223821],
    "splitmix64_seed_0_first": 0xE220A8397B1DCDAF,
    "pcg32_srandom_42_54": [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E],
}

if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "rng_kat.json"), "w") as f:
        json.dump(KAT, f, indent=1)
