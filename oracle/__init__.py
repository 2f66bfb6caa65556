"""CPU oracle of the SZSNN step (test infrastructure only; see pyoracle.py)."""
