"""CPU oracle for knn_aggregate_aniso -- test infrastructure only (see ``plain.py``)."""
