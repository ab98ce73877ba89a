"""Deck writer and the seeded problem generator."""

from exploringsycl_amd.deck import (Deck, config_deck, problem_seed,
                                    seeded_rectangles, splitmix64)


def test_splitmix64_known_output():
    # First outputs for seed 0 (published with the algorithm's reference
    # implementation, Vigna's splitmix64.c).
    g = splitmix64(0)
    assert next(g) == 0xE220A8397B1DCDAF
    assert next(g) == 0x6E789E6AA1B965F4


def test_rectangles_are_deterministic_and_in_range():
    a = seeded_rectangles(32, problem_seed(3, 7), density="log10",
                          density_range=(-1.0, 2.0))
    b = seeded_rectangles(32, problem_seed(3, 7), density="log10",
                          density_range=(-1.0, 2.0))
    assert a == b
    for r in a:
        assert 0 <= r["xmin"] <= r["xmax"] <= 10 and 0 <= r["ymin"] <= r["ymax"] <= 10
        assert 0.1 <= r["density"] <= 100 and 0.1 <= r["energy"] <= 25
    assert seeded_rectangles(4, 1) != seeded_rectangles(4, 2)


def test_config_decks_render():
    d = config_deck(2)
    text = d.render()
    assert "x_cells=4096" in text and "coefficient_inverse_density" in text
    assert text.count("geometry=rectangle") == 4 + 16
    assert "end_step=5" in text
    d1 = config_deck(1)
    assert "x_cells=512" in d1.render() and "coefficient_density" in d1.render()


def test_deck_round_trips_exact_values(tmp_path):
    d = Deck.with_random_rectangles(8, 8, 3, 99)
    p = d.write(tmp_path / "tea.in")
    text = open(p).read()
    for s in d.states[1:]:
        assert repr(s["xmin"]) in text and repr(s["density"]) in text


def test_write_deck_is_the_benchmark_field(tmp_path):
    """write_deck's decks are the time-to-solution field of bench.py (64
    seeded rectangles, log10 density in [-2, 2], K = 1/density, 1e-10) with
    each solver's keys, and the learned one comes with the network file."""
    from exploringsycl_amd import write_deck
    from exploringsycl_amd.deck import SOLVER_KEYS, tts_deck
    for solver in SOLVER_KEYS:
        d = tmp_path / solver
        assert write_deck.main([str(d), "--solver", solver, "--cells", "128",
                                "--max-iters", "50"]) == 0
        text = (d / "tea.in").read_text()
        assert text == tts_deck(solver, 128, 50).render()
        assert "rel_eps=1e-10" in text and "max_iters=50" in text.replace(" ", "")
        for key in SOLVER_KEYS[solver]:
            assert key in text
        assert (d / "net.bin").exists() == (solver == "fcg")
    assert text.count("geometry=rectangle") == 64 + 4


def test_counter_summary_groups_by_kernel_and_grid(tmp_path):
    """rocprof_summary counters: the mean per dispatch of each counter, per
    kernel and grid size, summing a counter's per-dimension rows within a
    dispatch."""
    import csv
    from exploringsycl_amd import rocprof_summary
    rows = [("void k1(int)", 256, 1, "SQ_WAVES", 4.0), ("void k1(int)", 256, 1, "SQ_WAVES", 6.0),
            ("void k1(int)", 256, 2, "SQ_WAVES", 20.0), ("void k1(int)", 512, 3, "SQ_WAVES", 7.0),
            ("void k2(float)", 64, 4, "GRBM_GUI_ACTIVE", 3.0)]
    with open(tmp_path / "p_1_counter_collection.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Grid_Size", "Dispatch_Id", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)
    got = rocprof_summary.all_counter_means(str(tmp_path))
    assert got[("k1", "256")] == {"SQ_WAVES": 15.0}
    assert got[("k1", "512")] == {"SQ_WAVES": 7.0}
    assert got[("k2", "64")] == {"GRBM_GUI_ACTIVE": 3.0}
