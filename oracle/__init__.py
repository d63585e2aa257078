"""The oracle: CPU restatement of the reference solver. Test infrastructure only."""
