"""Runs a TeaLeaf host program on a deck and returns its run record."""

import json
import os
import shutil
import subprocess
import tempfile

from .native import TEALEAF_BIN, TeaHipError

TEA_PROBLEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                            "data", "tea.problems")


def run_deck(deck, binary=TEALEAF_BIN, workdir=None, extra_args=(), env=None,
             timeout=600, problems=TEA_PROBLEMS):
    """Runs ``binary`` on ``deck`` (a Deck) in ``workdir`` (a fresh temporary
    directory by default).  Returns (record, stdout) where record is the
    parsed --json output."""
    if not os.path.exists(binary):
        raise TeaHipError(f"{binary} is not built")
    own = workdir is None
    workdir = workdir or tempfile.mkdtemp(prefix="tea_")
    os.makedirs(workdir, exist_ok=True)
    try:
        deck.write(os.path.join(workdir, "tea.in"))
        if problems and os.path.exists(problems):
            dst = os.path.join(workdir, "tea.problems")
            if os.path.lexists(dst):
                os.remove(dst)
            shutil.copyfile(problems, dst)
        cmd = [binary, "--json", "run.json", "--quiet"] + list(extra_args)
        e = dict(os.environ)
        if env:
            e.update(env)
        p = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True,
                           timeout=timeout, env=e)
        if p.returncode not in (0, 3):
            raise TeaHipError(f"{cmd} exited {p.returncode}:\n{p.stdout[-2000:]}"
                              f"\n{p.stderr[-2000:]}")
        with open(os.path.join(workdir, "run.json")) as f:
            record = json.load(f)
        return record, p.stdout
    finally:
        if own:
            shutil.rmtree(workdir, ignore_errors=True)
