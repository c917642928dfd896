"""Automata at benchmark scale for the front-end fixtures (regenerated from
seeded generators; tests/golden/frontend/large.json holds the SHA-256 of the
reference front-end's dump of each and its per-group counts)."""
import numpy as np

from ngap_amd import synth
from ngap_amd.groups import write_anml
from tests import nfa_helpers as H

# name: (--group-num, --max-nfa-size)
LARGE_CASES = {
    "regex20k": (1, -1),      # the regex20k workload's automaton, one group
    "batch64": (64, -1),      # 64 automata of the batch1024 generator (seed 1002)
    "wide3000": (7, -1),      # random automaton, 30 % of states with 5-8 successors
    "wide3000_max": (5, 200),  # the same with --max-nfa-size 200
}


def case_groups(name):
    """The groups whose ANML the case writes (None for wide cases)."""
    if name == "regex20k":
        return synth.regex20k()[0]
    if name == "batch64":
        return synth.batch_groups(np.random.default_rng(1002), 64)
    return None


def write_case(name, path):
    groups = case_groups(name)
    if groups is not None:
        synth.write_groups_anml(path, groups)
    else:
        write_anml(path, H.random_anml_states(7, n=3000, n_cc_hint=40, wide_p=0.3))
    return groups
