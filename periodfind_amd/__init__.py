"""MI355X-native batched periodogram engine with periodfind's API."""
