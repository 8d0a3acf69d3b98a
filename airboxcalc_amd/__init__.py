"""MI355X-native AirboxCalc compute core."""
