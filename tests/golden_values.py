"""The reference's quoted known answers, from tests/golden/reference_known_answers.json."""
import json
import os

_PATH = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


def known(name):
    with open(_PATH) as f:
        return json.load(f)[name]
