"""Writes tests/golden/decks/<deck>.json: the multigroup decks of the
reference (input_files/*.yaml: the LA-13511 one- and two-group cases and the
C5G7 assembly) read by abeille_amd.deck and flattened with Model.to_dict -
numbers, ids and region strings, not the deck text - so that the GPU tests
can run them where the reference is absent.

    python tests/golden/make_deck_fixtures.py [/root/reference/input_files]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from abeille_amd.deck import model_from_deck  # noqa: E402

DECKS = ["PUa-1-0-IN", "PUb-1-0-IN", "PUa-1-0-SL", "PUb-1-0-SL", "PUa-1-1-SL", "PUa-1-2-SL",
         "Ua-1-1-IN", "Ua-1-1-CY", "UD2O-2-1-SL", "c5g7"]


def main(src):
    out = os.path.join(HERE, "decks")
    os.makedirs(out, exist_ok=True)
    for name in DECKS:
        m = model_from_deck(os.path.join(src, name + ".yaml"))
        d = m.to_dict()
        d["deck"] = name + ".yaml"
        with open(os.path.join(out, name + ".json"), "w") as f:
            json.dump(d, f, separators=(",", ":"), sort_keys=True)
            f.write("\n")
        print(name, os.path.getsize(os.path.join(out, name + ".json")), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/input_files")
