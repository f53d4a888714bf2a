"""Parallel-tempering diagnostics of saved runs (CPU).  The cases and their
known answers follow the reference's tests/test_pt_analysis.py: weighted
acceptance, trip means over finished trips, skipped runs, the heatmap's
walker order, temperatures and irregular time edges, repeated sweeps, the
all-walkers title and the rejection of invalid histories."""

import copy
import json

import matplotlib
import numpy as np
import pytest

matplotlib.use("Agg")
import matplotlib.pyplot as plt  # noqa: E402

from utils.pt_analysis import PTAnalysis  # noqa: E402


def saved_run():
    return {"samples": {}, "metadata": {
        "L": 16, "temperatures": [0.8, 0.9, 1.0],
        "parallel_tempering": {
            "exchange_interval": 10, "exchange_seed": 9, "sweeps": 1000, "exchange_rounds": 100,
            "next_parity": 0, "diagnostics_scope": "since_wrapper_creation_including_thermalization",
            "attempts": [50, 50], "accepted": [40, 10], "acceptance_rates": [0.8, 0.2],
            "walker_to_temperature": [2, 0, 1], "temperature_to_walker": [1, 2, 0],
            "walker_round_trips": [2, 1, 0], "walker_round_trip_total_sweeps": [600, 400, 0],
            "mean_round_trip_sweeps": 1000 / 3}}}


@pytest.fixture(autouse=True)
def _close_figures():
    yield
    plt.close("all")


def test_rates_are_weighted_and_trip_means_cover_finished_trips():
    run = saved_run()
    run["metadata"]["parallel_tempering"].update(attempts=[50, 25], accepted=[40, 10])
    res = PTAnalysis({16: run}).summary()[16]
    assert res["aggregate_acceptance"] == 50 / 75
    assert res["acceptance_rates"] == [0.8, 0.4]
    assert res["walker_round_trips"] == [2, 1, 0]
    assert res["walker_mean_round_trip_sweeps"] == [300, 400, None]
    assert res["mean_round_trip_sweeps"] == pytest.approx(1000 / 3)
    assert res["diagnostics_scope"] == "since_wrapper_creation_including_thermalization"
    json.dumps(res, allow_nan=False)


def test_unattempted_edges_and_no_trips_give_no_means():
    run = saved_run()
    run["metadata"]["parallel_tempering"].update(attempts=[0, 0], accepted=[0, 0], walker_round_trips=[0, 0, 0],
                                                 walker_round_trip_total_sweeps=[0, 0, 0])
    res = PTAnalysis({16: run}).summary()[16]
    assert res["acceptance_rates"] == [None, None]
    assert res["aggregate_acceptance"] is None and res["mean_round_trip_sweeps"] is None


def test_runs_without_pt_are_skipped_and_plots_leave_inputs_alone():
    runs = {16: saved_run(), 32: {"metadata": {}}}
    before = copy.deepcopy(runs)
    analysis = PTAnalysis(runs)
    assert list(analysis.summary()) == [16]
    assert len(analysis.plot(show=False).axes) == 2
    assert runs == before
    assert "including thermalization" in analysis.format_summary()
    assert "exclude unfinished trips" in analysis.format_summary()
    empty = PTAnalysis({32: {"metadata": {}}})
    assert empty.summary() == {} and empty.plot(show=False) is None


def test_no_history_no_heatmap():
    analysis = PTAnalysis({16: saved_run()})
    assert analysis.plot_walkers(show=False) is None and analysis.walker_heatmap(16) is None
    assert not analysis.summary()[16]["history_available"]


def test_heatmap_orders_walkers_and_places_irregular_time_edges():
    run = saved_run()
    run["metadata"]["thermalization_sweeps"] = 10
    run["metadata"]["parallel_tempering"]["walker_history"] = {
        "walker_ids": [2, 0], "sweeps": [0, 10, 11], "temperature_indices": [[2, 0], [1, 1], [0, 2]]}
    before = copy.deepcopy(run)
    analysis = PTAnalysis({16: run})
    ids, edges, temps = analysis.walker_heatmap(16)
    np.testing.assert_array_equal(ids, [0, 2])
    np.testing.assert_array_equal(edges, [0, 5, 10.5, 11])
    np.testing.assert_array_equal(temps, [[0.8, 0.9, 1.0], [1.0, 0.9, 0.8]])
    ax = analysis.plot_walkers(show=False).axes[0]
    np.testing.assert_array_equal(ax.collections[0].get_array(), temps)
    np.testing.assert_array_equal(ax.collections[0].get_coordinates()[0, :, 0], edges)
    assert [t.get_text() for t in ax.get_yticklabels()] == ["0", "2"]
    assert run == before and analysis.summary()[16]["history_available"]


def test_heatmap_of_sixty_walkers():
    run = saved_run()
    T = np.linspace(0.36, 0.42, 60)
    run["metadata"]["temperatures"] = T
    positions = np.stack([np.arange(60), np.roll(np.arange(60), 1)])
    run["metadata"]["parallel_tempering"]["walker_history"] = {
        "walker_ids": np.arange(60), "sweeps": [0, 100], "temperature_indices": positions}
    ax = PTAnalysis({16: run}).plot_walkers(show=False).axes[0]
    assert "all walkers" in ax.get_title()
    np.testing.assert_array_equal(ax.collections[0].get_array(), T[positions].T)


def test_repeated_sweeps_keep_the_last_record_and_one_time_gets_unit_width():
    run = saved_run()
    run["metadata"]["parallel_tempering"]["walker_history"] = {
        "walker_ids": [0], "sweeps": [0, 0], "temperature_indices": [[0], [2]]}
    ids, edges, temps = PTAnalysis({16: run}).walker_heatmap(16)
    np.testing.assert_array_equal(temps, [[1.0]])
    np.testing.assert_array_equal(edges, [-0.5, 0.5])


@pytest.mark.parametrize("positions", [[[0], [3]], [[0.0], [1.0]], [[0, 1]]])
def test_invalid_histories_are_rejected(positions):
    run = saved_run()
    run["metadata"]["parallel_tempering"]["walker_history"] = {
        "walker_ids": [0], "sweeps": [0, 10], "temperature_indices": positions}
    with pytest.raises(ValueError):
        PTAnalysis({16: run})


def test_runs_of_one_size_under_several_models_keep_their_keys():
    analysis = PTAnalysis({16: saved_run(), (16, 1): saved_run(), 8: saved_run()})
    assert list(analysis.summary()) == [8, 16, (16, 1)]
