"""CPU oracle of the gample PIC step -- test infrastructure only (see pyoracle)."""
