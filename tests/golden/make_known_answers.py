"""Writes tests/golden/known_answers.json: published constants and test vectors
(no reference code is run; every value is a mathematical or standard-library
fact cited inline)."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

answers = {
    # C++ standard, [rand.predef]/3: 10000th consecutive invocation of a
    # default-constructed std::mt19937 produces 4123659995.
    "mt19937_10000th": 4123659995,
    # First three outputs of SplitMix64 (Steele, Lea & Flood 2014) seeded 0.
    "splitmix64_seed0": ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x06c45d188009454f"],
    # Madelung constant of the rock-salt (NaCl) lattice, referred to the
    # nearest-neighbour distance.
    "madelung_nacl": 1.747564594633182,
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(answers, f, indent=2)
        f.write("\n")
